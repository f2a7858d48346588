function cfdCorrectEquation(theEquationName, iComponent)
% Shadows uFVM/src/fvm/solve/correct/cfdCorrectEquation.m (same signature):
% U component iComponent, or for p the PP, velocity, pressure and mass-flux
% corrections, on the GPU case.  With Domain.ufvm.sync = 'fields' or 'all'
% a pressure correction writes U, p and mdot_f back to Domain.
global Domain
if nargin == 1
    iComponent = 1;
end
ufvm_mex('correct', Domain.ufvm.case, theEquationName, iComponent);
if strcmp(theEquationName, 'p') && ~strcmp(Domain.ufvm.sync, 'none')
    ufvmSyncFields({'U', 'p', 'mdot_f'});
end
end
