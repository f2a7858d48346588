function cfdUpdateFields
% Shadows uFVM/src/fvm/update/cfdUpdateFields.m: boundary values,
% gradients, face diffusivity and velocity scale of U and p
% (cfdUpdateEquations.m:14-24) on the GPU case made by ufvmSetup.
global Domain
ufvm_mex('update_fields', Domain.ufvm.case);
end
