/*
 * cuFoamPCG: CG with a none / diagonal / DIC / GAMG / GAMG32 preconditioner
 * (symmetric matrices; no counterpart in the reference) on the GPU through the
 * gis library, selected as "cuFoamPCG" in fvSolution.  Its interface is that
 * of the reference's wrappers (cuFoamCG/cuFoamCG.H:8-31).
 */
#ifndef CUFOAMPCG_H
#define CUFOAMPCG_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamPCG : public gisLduSolver {
public:
    TypeName("cuFoamPCG");

    cuFoamPCG(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamPCG() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMPCG_H
