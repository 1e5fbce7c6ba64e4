#include "gisLduSolver.H"

Foam::gisLduSolver::gisLduSolver(const word& fieldName, const lduMatrix& matrix,
                                 const FieldField<Field, scalar>& interfaceBouCoeffs,
                                 const FieldField<Field, scalar>& interfaceIntCoeffs,
                                 const lduInterfaceFieldPtrsList& interfaces,
                                 const dictionary& solverControls)
    : lduMatrix::solver(fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs, interfaces,
                        solverControls) {}

int Foam::gisLduSolver::preconditionerId() const {
    word name("none");
    if (controlDict_.isDict("preconditioner")) {
        name = controlDict_.subDict("preconditioner").lookupOrDefault<word>("preconditioner",
                                                                             "none");
    } else {
        name = controlDict_.lookupOrDefault<word>("preconditioner", "none");
    }
    if (name == "none") return GIS_PRECOND_NONE;
    if (name == "diagonal") return GIS_PRECOND_DIAGONAL;
    if (name == "GAMG") return GIS_PRECOND_GAMG;
    if (name == "GAMG32") return GIS_PRECOND_GAMG32;
    if (name == "DIC") return GIS_PRECOND_DIC;
    if (name == "DILU") return GIS_PRECOND_DILU;
    FatalErrorInFunction << "unknown preconditioner " << name
                         << " (valid: none diagonal DIC DILU GAMG GAMG32)" << abort(FatalError);
    return GIS_PRECOND_NONE;
}

Foam::solverPerformance Foam::gisLduSolver::solveWith(const word& solverName,
                                                      gisFoam::Family family, int id,
                                                      scalarField& psi,
                                                      const scalarField& source) const {
    static_assert(sizeof(label) == sizeof(int), "gis takes 32-bit labels (WM_LABEL_SIZE=32)");
    static_assert(sizeof(scalar) == sizeof(double), "gis takes double scalars (WM_DP)");

    const lduAddressing& addr = matrix().lduAddr();
    const gisFoam::Addressing a{&addr, addr.size(), addr.lowerAddr().size(),
                                addr.lowerAddr().cdata(), addr.upperAddr().cdata()};
    // a symmetric lduMatrix holds one triangle: lower() then returns upper()
    const bool faces = !matrix().diagonal();
    const gisFoam::Coefficients c{matrix().diag().cdata(),
                                  faces ? matrix().lower().cdata() : nullptr,
                                  faces ? matrix().upper().cdata() : nullptr};

    gisFoam::Controls ctl;
    ctl.tolerance = tolerance_;
    ctl.maxIter = maxIter_;
    ctl.relTol = relTol_;
    ctl.minIter = minIter_;
    const word convergence = controlDict_.lookupOrDefault<word>("convergence", "reference");
    if (convergence != "reference" && convergence != "openfoam") {
        FatalErrorInFunction << solverName << ": convergence must be reference or openfoam, not "
                             << convergence << abort(FatalError);
    }
    ctl.openfoam = (convergence == "openfoam");
    ctl.device = controlDict_.lookupOrDefault<label>("device", 0);

    const gisFoam::Outcome o = gisFoam::solveAs(family, id, a, c, psi.data(), source.cdata(), ctl);
    if (o.fatal) {
        FatalErrorInFunction << solverName << " for " << fieldName() << ": " << o.message
                             << abort(FatalError);
    }
    if (o.warn) {
        WarningInFunction << solverName << " for " << fieldName() << ": " << o.message << endl;
    }

    solverPerformance perf(solverName, fieldName());
    perf.initialResidual() = o.perf.initial_residual;
    perf.finalResidual() = o.perf.final_residual;
    perf.nIterations() = o.perf.n_iterations;
    return perf;
}
