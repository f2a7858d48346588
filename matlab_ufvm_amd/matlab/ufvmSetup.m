function ufvmSetup(varargin)
% ufvmSetup  Move uFVM's SIMPLE path onto the GPU (libufvm through ufvm_mex).
%
%   ufvmSetup                     after the reference's case set-up (mesh
%   ufvmSetup('pSolver', 'pcg')   processed, U and p defined, fields
%                                 initialised), before cfdRunCase
%
% Gathers `global Domain` -- theMesh (cfdProcessOpenFoamMesh.m), the U and p
% models' boundary conditions (theEquation.bcs), relaxation factors,
% nCorrectors, the reference cell (Domain.iFixedElement, cfdReadSystem.m:
% 238-251) and the laplacian coefficient / density fields -- into plain
% arrays, creates the library case, and copies the current U, p and mdot_f
% into it.  The .m shims in shims/case then route cfdUpdateFields,
% cfdTransientUpdate, cfdAssembleEquation, cfdSolveEquation,
% cfdCorrectEquation and cfdAssembleAndCorrectEquation to it.
%
% Options (name/value):
%   'pSolver'    'amg' (the reference's cfdApplyAMG, default), 'pcg' (the
%                FP32-preconditioned AMG-PCG), 'pcg-frozen' (PCG keeping the
%                first hierarchy)
%   'pcgTol'     relative residual of the PCG (1e-8); also 'pcgPrecision',
%                'pcgSweeps', 'pcgMaxCoarseCells', 'pcgSmoother'
%   'uSmoother'  'ILU' (the reference's DILU, default) or 'MCGS'
%   'sync'       'fields' (default: U, p and mdot_f written back to Domain
%                after each pressure correction, so cfdWriteResults and the
%                residual plots see them), 'all' (also Domain.coefficients{1}
%                after every assembly), 'none' (only ufvmSyncFields)
%   'device'     GPU index (0)
global Domain

opts = struct('pSolver', 'amg', 'sync', 'fields', 'device', 0);
for k = 1:2:numel(varargin)
    opts.(varargin{k}) = varargin{k + 1};
end
ufvm_mex('init', opts.device);

theMesh = cfdGetMesh;
N = theMesh.numberOfElements;
F = theMesh.numberOfFaces;
Fi = theMesh.numberOfInteriorFaces;
faces = theMesh.faces;
elements = theMesh.elements;

m = struct();
m.nElements = N;
m.nFaces = F;
m.nInteriorFaces = Fi;
m.iOwner = int32([faces.iOwner]');
m.iNeighbour = int32([faces(1:Fi).iNeighbour]');
m.Sf = reshape([faces.Sf], 3, F)';
m.area = [faces.area]';
m.CN = reshape([faces.CN], 3, F)';
m.eCN = reshape([faces.eCN], 3, F)';
m.gDiff = [faces.gDiff]';
m.T = reshape([faces.T], 3, F)';
m.gf = [faces.gf]';
m.faceCentroid = reshape([faces.centroid], 3, F)';
m.volume = [elements.volume]';
m.elementCentroid = reshape([elements.centroid], 3, N)';
m.csize = double(theMesh.csize(:));
m.cconn = ufvmFlatten(theMesh.cconn);
m.iOwnerNeighbourCoef = [faces(1:Fi).iOwnerNeighbourCoef]';
m.iNeighbourOwnerCoef = [faces(1:Fi).iNeighbourOwnerCoef]';
m.patchStartFace = [theMesh.boundaries.startFace]';
m.patchNFaces = [theMesh.boundaries.numberOfBFaces]';
m.patchType = {theMesh.boundaries.type};

% boundary conditions: theEquation.bcs{iPatch}.type / .value (the value a
% string the reference evaluates, e.g. '[1;0;0]')
U = cfdGetModel('U');
p = cfdGetModel('p');
P = numel(theMesh.boundaries);
o = struct();
o.uBC = cell(1, P);
o.pBC = cell(1, P);
o.uValue = zeros(P, 3);
o.pValue = zeros(P, 1);
for iPatch = 1:P
    o.uBC{iPatch} = U.bcs{iPatch}.type;
    o.pBC{iPatch} = p.bcs{iPatch}.type;
    if strcmp(U.bcs{iPatch}.type, 'fixedValue')
        v = eval(U.bcs{iPatch}.value);
        if numel(v) ~= 3
            error('ufvm:setup', 'patch %d: only uniform fixedValue U is supported', iPatch);
        end
        o.uValue(iPatch, :) = reshape(v, 1, 3);
    end
    if strcmp(p.bcs{iPatch}.type, 'fixedValue')
        v = eval(p.bcs{iPatch}.value);
        if numel(v) ~= 1
            error('ufvm:setup', 'patch %d: only uniform fixedValue p is supported', iPatch);
        end
        o.pValue(iPatch) = v;
    end
end

% relaxation, correctors, reference cell (cfdDefineEquation.m:84-99)
o.urfU = U.urf;
o.urfP = p.urf;
algorithm = cfdGetAlgorithm;
o.nCorrectors = Domain.foam.fvSolution.(algorithm).nCorrectors;
if isfield(Domain, 'iFixedElement') && ~isempty(Domain.iFixedElement)
    o.pRefCell = Domain.iFixedElement;
else
    o.pRefCell = 0;
end

% the laplacian's coefficient and the density: constant fields on this path
for iTerm = 1:numel(U.terms)
    theTerm = U.terms{iTerm};
    if strcmp(theTerm.name, 'Diffusion')
        nu = cfdGetMeshField(theTerm.coefficientName);
        if any(nu.phi(:) ~= nu.phi(1))
            error('ufvm:setup', 'the laplacian coefficient must be constant');
        end
        o.nu = nu.phi(1);
    elseif strcmp(theTerm.name, 'Transient')
        o.transient = 1;
        time = cfdGetTime;
        o.dt = time.deltaT;
    elseif strcmp(theTerm.name, 'Source') && contains(theTerm.expression, 'transp')
        o.stressTerm = 1;
    end
end
rho = cfdGetMeshField('rho_Ueq');
if ~isempty(rho)
    if any(rho.phi(:) ~= rho.phi(1))
        error('ufvm:setup', 'the density must be constant');
    end
    o.rho = rho.phi(1);
end
o.smootherType = p.theSolutionMethods.smootherType;
% momentum solves (cfdSolver.m:11-12: 10 sweeps of U's smoother); the path
% has the reference's DILU ('ILU') or multicolour GS ('MCGS', an option)
o.uSmoother = U.theSolutionMethods.smootherType;
mg = p.multigrid;
for f = {'cycleType', 'maxCycles', 'termination', 'maxCoarseLevels', 'preSweep', 'postSweep'}
    o.(f{1}) = mg.(f{1});
end
for f = {'pSolver', 'uSmoother', 'pcgTol', 'pcgPrecision', 'pcgSweeps', 'pcgMaxCoarseCells', ...
         'pcgSmoother'}
    if isfield(opts, f{1})
        o.(f{1}) = opts.(f{1});
    end
end

id = ufvm_mex('case_create', m, o);

% the state the reference set up (possibly a restart): U, p, mdot_f
Uf = cfdGetMeshField('U');
ufvm_mex('set_field', id, 'U', Uf.phi);
pf = cfdGetMeshField('p');
ufvm_mex('set_field', id, 'p', pf.phi);
mf = cfdGetMeshField('mdot_f', 'Faces');
ufvm_mex('set_field', id, 'mdot_f', mf.phi);

Domain.ufvm = struct('case', id, 'sync', opts.sync, 'transient', isfield(o, 'transient'));
end
