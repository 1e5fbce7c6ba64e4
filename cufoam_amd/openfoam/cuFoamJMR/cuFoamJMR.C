#include "cuFoamJMR.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamJMR, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamJMR>
    addcuFoamJMRSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamJMR>
    addcuFoamJMRAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamJMR::cuFoamJMR(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamJMR::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_JMR, psi, source);
}
