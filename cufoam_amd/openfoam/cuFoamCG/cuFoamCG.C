#include "cuFoamCG.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamCG, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamCG>
    addcuFoamCGSymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamCG::cuFoamCG(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamCG::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_CG, psi, source);
}
