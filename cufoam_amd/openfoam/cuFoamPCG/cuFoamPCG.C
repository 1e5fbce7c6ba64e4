#include "cuFoamPCG.H"

namespace Foam {

defineTypeNameAndDebug(cuFoamPCG, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<cuFoamPCG>
    addcuFoamPCGSymMatrixConstructorToTable_;

}  // namespace Foam

Foam::cuFoamPCG::cuFoamPCG(const word& fieldName, const lduMatrix& matrix,
                   const FieldField<Field, scalar>& interfaceBouCoeffs,
                   const FieldField<Field, scalar>& interfaceIntCoeffs,
                   const lduInterfaceFieldPtrsList& interfaces,
                   const dictionary& solverControls)
    : gisLduSolver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                   solverControls) {}

Foam::solverPerformance Foam::cuFoamPCG::solve(scalarField& psi, const scalarField& source,
                                           const direction) const {
    return solveWith(typeName, gisFoam::Family::PCG, preconditionerId(), psi, source);
}
