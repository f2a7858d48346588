function [initialResidual, finalResidual] = cfdSolveEquation(theEquationName)
% Shadows uFVM/src/fvm/solve/cfdSolveEquation.m (same signature): U by
% 10 sweeps of its smoother (cfdSolver.m), p by cfdApplyAMG or, as ufvmSetup
% chose, the FP32-preconditioned AMG-PCG, on the GPU case.  As cfdApplyAMG
% does, the cycle counter goes to theEquation.multigrid.nCycles.
global Domain
[initialResidual, finalResidual, nCycles] = ufvm_mex('solve', Domain.ufvm.case, theEquationName);
theEquation = cfdGetModel(theEquationName);
if theEquation.multigrid.isActive
    theEquation.multigrid.nCycles = num2str(nCycles);
    cfdSetModel(theEquation);
end
end
