function [rmsResidual, maxResidual, initialResidual, finalResidual] = cfdAssembleAndCorrectEquation(theEquationName)
% Shadows uFVM/src/fvm/assemble/cfdAssembleAndCorrectEquation.m (same
% signature and loop: per component, nCorrectors passes of assemble / solve
% / correct for p, one for U), each step through the shims beside it.
global Domain
theEquation = cfdGetModel(theEquationName);
if strcmp(theEquation.type, 'Scalar')
    theNumberOfComponents = 1;
else
    theNumberOfComponents = 3;
end
if strcmp(theEquationName, 'p')
    algorithm = cfdGetAlgorithm;
    nCorrectors = Domain.foam.fvSolution.(algorithm).nCorrectors;
else
    nCorrectors = 1;
end
rmsResidual = zeros(1, theNumberOfComponents);
maxResidual = zeros(1, theNumberOfComponents);
initialResidual = zeros(1, theNumberOfComponents);
finalResidual = zeros(1, theNumberOfComponents);
for iComponent = 1:theNumberOfComponents
    for iCorrector = 1:nCorrectors
        [rmsResidual(iComponent), maxResidual(iComponent)] = cfdAssembleEquation(theEquationName, iComponent);
        [initialResidual(iComponent), finalResidual(iComponent)] = cfdSolveEquation(theEquationName);
        cfdCorrectEquation(theEquationName, iComponent);
    end
    cfdStoreResiduals(theEquationName, iComponent, rmsResidual(iComponent));
end
end
