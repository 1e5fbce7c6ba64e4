#include "cuFoamMR.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamMR, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamMR>
    addcuFoamMRSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamMR>
    addcuFoamMRAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamMR::cuFoamMR(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamMR::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_MR, psi, source);
}
