/*
 * cuFoamJMR: minimal residual with symmetric Jacobi scaling (PreconditionedMRJacobi) on the GPU through the gis library, selected as
 * "cuFoamJMR" in fvSolution (the interface of the reference's cuFoamJMR/cuFoamJMR.H).
 */
#ifndef CUFOAMJMR_H
#define CUFOAMJMR_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamJMR : public gisLduSolver {
public:
    TypeName("cuFoamJMR");

    cuFoamJMR(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamJMR() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMJMR_H
