function residual = cfdComputeResidualsArray(theCoefficients)
% Shadows uFVM/src/fvm/compute/cfdComputeResidualsArray.m (same signature):
% r = bc - ac .* dphi - sum(anb .* dphi(cconn)) on the GPU, each row summed
% in cconn order (the same bits as the reference's loop).
sys = ufvmSystemFor(theCoefficients);
residual = reshape(ufvm_mex('residuals', sys, 1), size(theCoefficients.ac));
end
