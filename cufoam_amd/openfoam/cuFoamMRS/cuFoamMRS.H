/*
 * cuFoamMRS: minimal residual with invM = D^-1 applied as an SpMV on the GPU through the gis library, selected as
 * "cuFoamMRS" in fvSolution (the interface of the reference's cuFoamMRS/cuFoamMRS.H).
 */
#ifndef CUFOAMMRS_H
#define CUFOAMMRS_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamMRS : public gisLduSolver {
public:
    TypeName("cuFoamMRS");

    cuFoamMRS(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamMRS() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMMRS_H
