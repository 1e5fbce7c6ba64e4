/*
 * cuFoamCG: unpreconditioned conjugate gradients (symmetric matrices) on the GPU through the gis library, selected as
 * "cuFoamCG" in fvSolution (the interface of the reference's cuFoamCG/cuFoamCG.H).
 */
#ifndef CUFOAMCG_H
#define CUFOAMCG_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamCG : public gisLduSolver {
public:
    TypeName("cuFoamCG");

    cuFoamCG(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamCG() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMCG_H
