#include "cuFoamJacobi.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamJacobi, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamJacobi>
    addcuFoamJacobiSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamJacobi>
    addcuFoamJacobiAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamJacobi::cuFoamJacobi(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamJacobi::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_JACOBI, psi, source);
}
