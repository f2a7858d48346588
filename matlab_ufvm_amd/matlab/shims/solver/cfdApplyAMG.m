function [initialResidual, finalResidual] = cfdApplyAMG(theEquation)
% Shadows uFVM/src/fvm/solve/AMG/cfdApplyAMG.m (same signature): the
% agglomeration of Domain.coefficients{1} and the V/W/F cycles on the GPU
% (parents, coarse operators and dphi bit-exact with the reference).  As the
% reference, it writes Domain.coefficients{1}.dphi, the coarse levels
% Domain.coefficients{2..maxLevels} and theEquation.multigrid.nCycles.
global Domain
smootherType = theEquation.theSolutionMethods.smootherType;
sys = ufvmSystemFor(cfdGetCoefficients(1));
[initialResidual, finalResidual, nCycle] = ufvm_mex('apply_amg', sys, theEquation.multigrid, ...
                                                    smootherType);
theCoefficients = cfdGetCoefficients(1);
theCoefficients.dphi = reshape(ufvm_mex('get_dphi', sys, 1), size(theCoefficients.dphi));
% the hierarchy cfdAgglomerate leaves behind: parents of every level but the
% coarsest, the coarse levels' coefficient structs (skipped when
% Domain.ufvm.skipCoarseLevels is true: nothing on the path reads them)
if ~isfield(Domain, 'ufvm') || ~isfield(Domain.ufvm, 'skipCoarseLevels') || ~Domain.ufvm.skipCoarseLevels
    nLevels = ufvm_mex('num_levels', sys);
    if nLevels > 1
        c1 = ufvm_mex('get_coefficients', sys, 1);
        theCoefficients.parents = c1.parents;
    end
    for iLevel = 2:nLevels
        cfdSetCoefficients(ufvm_mex('get_coefficients', sys, iLevel), iLevel);
    end
end
cfdSetCoefficients(theCoefficients, 1);
theEquation.multigrid.nCycles = num2str(nCycle);
cfdSetModel(theEquation);
end
