/*
 * cuFoamMRJacobi: minimal residual on the Jacobi-preconditioned residual on the GPU through the gis library, selected as
 * "cuFoamMRJacobi" in fvSolution (the interface of the reference's cuFoamMRJacobi/cuFoamMRJacobi.H).
 */
#ifndef CUFOAMMRJACOBI_H
#define CUFOAMMRJACOBI_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamMRJacobi : public gisLduSolver {
public:
    TypeName("cuFoamMRJacobi");

    cuFoamMRJacobi(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamMRJacobi() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMMRJACOBI_H
