/*
 * cuFoamRichardson: Richardson iteration, omega from 30 power steps (symmetric matrices) on the GPU through the gis library, selected as
 * "cuFoamRichardson" in fvSolution (the interface of the reference's cuFoamRichardson/cuFoamRichardson.H).
 */
#ifndef CUFOAMRICHARDSON_H
#define CUFOAMRICHARDSON_H

#include "gisLduSolver.H"

namespace Foam {

class cuFoamRichardson : public gisLduSolver {
public:
    TypeName("cuFoamRichardson");

    cuFoamRichardson(const word& fieldName, const lduMatrix& matrix,
          const FieldField<Field, scalar>& interfaceBouCoeffs,
          const FieldField<Field, scalar>& interfaceIntCoeffs,
          const lduInterfaceFieldPtrsList& interfaces, const dictionary& solverControls);

    virtual ~cuFoamRichardson() = default;

    virtual solverPerformance solve(scalarField& psi, const scalarField& source,
                                    const direction cmpt = 0) const;
};

}  // namespace Foam

#endif  // CUFOAMRICHARDSON_H
