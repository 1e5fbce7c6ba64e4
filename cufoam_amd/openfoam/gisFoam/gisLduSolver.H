/*
 * gisLduSolver -- the common part of the cuFoam<Name> lduMatrix::solver
 * wrappers: OpenFOAM's solver constructor and one solve sequence over the gis
 * library (gisFoamBridge.H).  Each wrapper adds its TypeName, its runtime
 * selection entries and the gis method it runs (cuFoamCG/cuFoamCG.H:8-31 and
 * siblings in the reference, which repeat the whole sequence per wrapper).
 *
 * fvSolution keys read besides OpenFOAM's own (tolerance, relTol, maxIter,
 * minIter):
 *   convergence  reference | openfoam  (default reference: tolerance and
 *                maxIter only, as the reference wrappers; openfoam: relTol
 *                and minIter too, through gis_foam_solve_controls)
 *   device       GPU index (default 0)
 *   preconditioner  none | diagonal | DIC | GAMG | GAMG32  (cuFoamPCG);
 *                   none | diagonal | DILU  (cuFoamPBiCG)
 */
#ifndef GIS_LDU_SOLVER_H
#define GIS_LDU_SOLVER_H

#include "lduMatrix.H"

#include "gisFoamBridge.H"

namespace Foam {

class gisLduSolver : public lduMatrix::solver {
protected:
    gisLduSolver(const word& fieldName, const lduMatrix& matrix,
                 const FieldField<Field, scalar>& interfaceBouCoeffs,
                 const FieldField<Field, scalar>& interfaceIntCoeffs,
                 const lduInterfaceFieldPtrsList& interfaces,
                 const dictionary& solverControls);

    // psi is updated in place; the performance carries the reference's
    // unnormalised initial and final L2 residuals and the iteration count.
    // family Plain: id is a GIS_<method>; PCG / PBiCG: a GIS_PRECOND_<name>.
    solverPerformance solveWith(const word& solverName, gisFoam::Family family, int id,
                                scalarField& psi, const scalarField& source) const;

    // the plain methods of the reference's wrappers
    solverPerformance solveWith(const word& solverName, int method, scalarField& psi,
                                const scalarField& source) const {
        return solveWith(solverName, gisFoam::Family::Plain, method, psi, source);
    }

    // the fvSolution preconditioner name as a GIS_PRECOND_ id (cuFoamPCG, cuFoamPBiCG)
    int preconditionerId() const;
};

}  // namespace Foam

#endif  // GIS_LDU_SOLVER_H
