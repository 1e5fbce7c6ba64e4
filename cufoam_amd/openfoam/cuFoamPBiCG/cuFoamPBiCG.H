/*
 * cuFoamPBiCG: BiCG with a none / diagonal / DILU preconditioner (symmetric and
 * asymmetric matrices; no counterpart in the reference) on the GPU through the
 * gis library, selected as "cuFoamPBiCG" in fvSolution.  Its interface is that
 * of the reference's wrappers (cuFoamBiCG/cuFoamBiCG.H:8-31).
 */
#ifndef CUFOAMPBICG_H
#define CUFOAMPBICG_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamPBiCG : public gisLduSolver {
public:
    TypeName("cuFoamPBiCG");

    cuFoamPBiCG(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamPBiCG() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMPBICG_H
