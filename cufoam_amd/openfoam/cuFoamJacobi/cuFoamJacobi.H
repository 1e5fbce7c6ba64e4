/*
 * cuFoamJacobi: residual-update Jacobi iteration on the GPU through the gis library, selected as
 * "cuFoamJacobi" in fvSolution (the interface of the reference's cuFoamJacobi/cuFoamJacobi.H).
 */
#ifndef CUFOAMJACOBI_H
#define CUFOAMJACOBI_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamJacobi : public gisLduSolver {
public:
    TypeName("cuFoamJacobi");

    cuFoamJacobi(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamJacobi() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMJACOBI_H
