#include "cuFoamBiCG.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamBiCG, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamBiCG>
    addcuFoamBiCGSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamBiCG>
    addcuFoamBiCGAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamBiCG::cuFoamBiCG(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamBiCG::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, GIS_BICG, psi, source);
}
