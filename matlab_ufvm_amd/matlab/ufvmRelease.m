function ufvmRelease
% ufvmRelease  Free the GPU case and the library context (after cfdRunCase).
global Domain
if isfield(Domain, 'ufvm')
    Domain = rmfield(Domain, 'ufvm');
end
ufvm_mex('release');
end
