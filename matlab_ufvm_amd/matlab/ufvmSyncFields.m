function ufvmSyncFields(names)
% ufvmSyncFields  Copy fields of the GPU case back into `global Domain`.
%
%   ufvmSyncFields                 U, p and mdot_f
%   ufvmSyncFields({'U', 'DU'})    the named fields
%
% Each field replaces the phi of the reference's mesh field of that name
% (cfdGetMeshField / cfdSetMeshField.m; mdot_f in the 'Faces' locale), in
% the reference's layout: U (N+Fb) x 3, p (N+Fb) x 1, mdot_f F x 1.
global Domain
if nargin < 1
    names = {'U', 'p', 'mdot_f'};
end
id = Domain.ufvm.case;
for k = 1:numel(names)
    name = names{k};
    if strcmp(name, 'mdot_f')
        theField = cfdGetMeshField(name, 'Faces');
    else
        theField = cfdGetMeshField(name);
    end
    if isempty(theField)
        error('ufvm:sync', 'Domain has no mesh field %s', name);
    end
    theField.phi = ufvm_mex('get_field', id, name);
    cfdSetMeshField(theField);
end
end
