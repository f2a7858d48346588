function sys = ufvmSystemFor(theCoefficients)
% ufvmSystemFor  The library system holding a coefficient struct of the
% reference (cfdSetupCoefficients.m: ac, anb, bc, dphi, cconn, csize).
%
% One system per distinct pattern (cconn) is kept for the session; each
% call uploads the struct's ac, anb, bc and dphi into it.
persistent cache
if isempty(cache)
    cache = struct('csize', {}, 'cconn', {}, 'sys', {});
end
csize = double(theCoefficients.csize(:));
cconn = ufvmFlatten(theCoefficients.cconn);
sys = [];
for k = 1:numel(cache)
    if isequal(cache(k).csize, csize) && isequal(cache(k).cconn, cconn)
        sys = cache(k).sys;
        break
    end
end
if isempty(sys)
    sys = ufvm_mex('system_create', csize, cconn);
    cache(end + 1) = struct('csize', csize, 'cconn', cconn, 'sys', sys);
end
ufvm_mex('set_coefficients', sys, double(theCoefficients.ac(:)), ...
         ufvmFlatten(theCoefficients.anb), double(theCoefficients.bc(:)));
ufvm_mex('set_dphi', sys, 1, double(theCoefficients.dphi(:)));
end
