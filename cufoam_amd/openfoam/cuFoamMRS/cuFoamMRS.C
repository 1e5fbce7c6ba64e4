#include "cuFoamMRS.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamMRS, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamMRS>
    addcuFoamMRSSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamMRS>
    addcuFoamMRSAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamMRS::cuFoamMRS(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamMRS::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_MRS, psi, source);
}
