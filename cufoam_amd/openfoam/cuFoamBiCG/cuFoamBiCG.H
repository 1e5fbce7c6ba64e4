/*
 * cuFoamBiCG: BiCG with the transpose built on the device (breakdown bound 1e-10) on the GPU through the gis library, selected as
 * "cuFoamBiCG" in fvSolution (the interface of the reference's cuFoamBiCG/cuFoamBiCG.H).
 */
#ifndef CUFOAMBICG_H
#define CUFOAMBICG_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamBiCG : public gisLduSolver {
public:
    TypeName("cuFoamBiCG");

    cuFoamBiCG(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamBiCG() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMBICG_H
