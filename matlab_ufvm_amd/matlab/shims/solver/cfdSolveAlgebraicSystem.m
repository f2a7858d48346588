function [initialResidual, finalResidual] = cfdSolveAlgebraicSystem(gridLevel, smootherType, numberOfIterations, rrf)
% Shadows uFVM/src/fvm/solve/cfdSolveAlgebraicSystem.m (same signature):
% the 'ILU' (DILU) or 'SOR' sweeps of Domain.coefficients{gridLevel} on the
% GPU in the reference's orderings, with its mean-|r| early exit; dphi is
% written back as the reference writes it.
if nargin == 3
    rrf = 0.1;
end
theCoefficients = cfdGetCoefficients(gridLevel);
sys = ufvmSystemFor(theCoefficients);
[initialResidual, finalResidual] = ufvm_mex('solve_algebraic_system', sys, 1, smootherType, ...
                                            numberOfIterations, rrf);
theCoefficients.dphi = reshape(ufvm_mex('get_dphi', sys, 1), size(theCoefficients.dphi));
cfdSetCoefficients(theCoefficients, gridLevel);
end
