function [rmsResidual, maxResidual] = cfdAssembleEquation(theEquationName, iComponent)
% Shadows uFVM/src/fvm/assemble/cfdAssembleEquation.m (same signature): the
% equation's coefficients assembled on the GPU case (U: momentum component
% iComponent; p: Rhie-Chow fluxes and the pressure-correction matrix).
% With Domain.ufvm.sync = 'all' the coefficients are also written to
% Domain.coefficients{1}, as the reference leaves them.
global Domain
if nargin == 1
    iComponent = 1;
end
[rmsResidual, maxResidual] = ufvm_mex('assemble', Domain.ufvm.case, theEquationName, iComponent);
if strcmp(Domain.ufvm.sync, 'all')
    sys = ufvm_mex('case_system', Domain.ufvm.case, theEquationName);
    cfdSetCoefficients(ufvm_mex('get_coefficients', sys, 1));
end
end
