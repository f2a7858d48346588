function cfdTransientUpdate
% Shadows uFVM/src/fvm/update/cfdTransientUpdate.m: U and mdot_f become the
% previous time step's fields of the GPU case.  The steady loop calls it
% too; there the path reads no previous-step field, so it does nothing.
global Domain
if Domain.ufvm.transient
    ufvm_mex('transient_update', Domain.ufvm.case);
end
end
