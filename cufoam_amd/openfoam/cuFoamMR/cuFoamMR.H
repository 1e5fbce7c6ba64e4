/*
 * cuFoamMR: the minimal residual iteration on the GPU through the gis library, selected as
 * "cuFoamMR" in fvSolution (the interface of the reference's cuFoamMR/cuFoamMR.H).
 */
#ifndef CUFOAMMR_H
#define CUFOAMMR_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamMR : public gisLduSolver {
public:
    TypeName("cuFoamMR");

    cuFoamMR(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamMR() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMMR_H
