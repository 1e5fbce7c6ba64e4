#include "cuFoamRichardson.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamRichardson, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamRichardson>
    addcuFoamRichardsonSymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamRichardson::cuFoamRichardson(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamRichardson::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_RICHARDSON, psi, source);
}
