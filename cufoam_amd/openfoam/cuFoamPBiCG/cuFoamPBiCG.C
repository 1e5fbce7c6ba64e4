#include "cuFoamPBiCG.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamPBiCG, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamPBiCG>
    addcuFoamPBiCGSymMatrixConstructorToTable_;

lduMatrix::solver::addasymMatrixConstructorToTable<cuFoamPBiCG>
    addcuFoamPBiCGAsymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamPBiCG::cuFoamPBiCG(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamPBiCG::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, gisFoam::Family::PBiCG, preconditionerId(), psi, source);
}
