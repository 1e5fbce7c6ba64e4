#include "cuFoamMRJacobi.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamMRJacobi, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamMRJacobi>
    addcuFoamMRJacobiSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamMRJacobi>
    addcuFoamMRJacobiAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamMRJacobi::cuFoamMRJacobi(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamMRJacobi::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_MRJACOBI, psi, source);
}
