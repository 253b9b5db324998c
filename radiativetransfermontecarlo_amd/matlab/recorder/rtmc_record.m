function rec = rtmc_record(refdir, name, N, outfile)
%RTMC_RECORD  Records the reference's own random draws for librtmc's replay mode.
%
%   rec = rtmc_record(refdir, name, N, outfile) runs one packet of N particles
%   of case `name` through the reference's functions (refdir: a checkout of
%   RadiativeTransferMonteCarlo) in the order of radiativeTransfer's serial
%   loop (radiativeTransfer.m:109-131): initializeParticle, then for every
%   observation time propagateParticleSmallDt (propagateParticle when
%   material.timeSteps = 1) and observeTime.  Every draw the reference makes
%   is stored per particle in librtmc's replay layout, replay(step, slot,
%   particle), with the slot tables of rtmc_init_particles,
%   rtmc_propagate_smalldt and rtmc_propagate_events (include/rtmc.h), next
%   to MATLAB's particle states after every call, its histogram E, T - mean(P.t)
%   of every sub-stepped call and the prepared material's tables, and saved to
%   outfile (MAT v7).  tests/test_matlab_replay.py replays such a file
%   through librtmc: bins and counts exactly, particle states to 1e-12.
%
%   The draws are caught by shadows of rand, randn, randi and poissrnd (the
%   shadows subfolder), which this function puts on the path for the run only;
%   each site is found by its source text, so a changed reference line stops
%   the recorder with an error instead of recording the wrong site.  Put this
%   folder (not shadows) on the path, or run from it.
%
%   Cases: 'readme3d' (3-D acoustic, spherical frame, with Q), 'elastic2d'
%   (2-D elastic, Q), 'halfspace3d' (3-D elastic under a free surface z = 0:
%   MaterialInterface's draws) and 'events3d' (3-D acoustic, event-driven).
%   Limits: at most one elastic plane boundary (MaterialInterface does not
%   know which boundary calls it; its draws are stored as boundary 1).
%   Requires the Statistics and Machine Learning Toolbox (poissrnd).
%
%   Example (from this folder):
%       rtmc_record('/path/to/RadiativeTransferMonteCarlo', 'readme3d', 2000, ...
%                   fullfile(repo, 'tests', 'golden', 'matlab', 'readme3d.mat'));

global RTMC_REC
shadows = fullfile(fileparts(mfilename('fullpath')), 'shadows');
addpath(refdir);
% the toolbox poissrnd, taken before the shadow goes on the path
p = which('poissrnd', '-all');
p = p(~startsWith(p, shadows));
if isempty(p)
    error('rtmc_record:toolbox', 'poissrnd (Statistics and Machine Learning Toolbox) is required');
end
old = cd(fileparts(p{1}));
hpois = @poissrnd;
cd(old);

[geometry, source, material, observation] = rtmc_case(name, N);
d = geometry.dimension;
acoustics = material.acoustics;
[~, E, bins, ibins, vals, Nt, t] = initializeObservation(geometry, acoustics, observation, N);
material = MaterialClass.prepareSigma(material, d);
nbnd = 0;
if isfield(geometry, 'bnd')
    nbnd = numel(geometry.bnd);
end
dt = mean(diff(t));

RTMC_REC = struct('on', false, 'poissrnd', hpois, 'sites', rtmc_sites(refdir), 'call', []);
addpath(shadows, '-begin');
cleanup = onCleanup(@() rtmc_stop(shadows));
RTMC_REC.on = true;

rec_begin('init', 8, N);
P = initializeParticle(N, d, acoustics, source);
init_draws = rec_end(1);
init_state = state(P);

draws = cell(1, Nt - 1);
states = cell(1, Nt - 1);
teff = nan(1, Nt - 1);
tval = nan(1, Nt - 1);
for it = 2:Nt
    t_val = t(1) + (it - 1)*dt;                     % radiativeTransfer.m:117
    tval(it - 1) = t_val;
    if material.timeSteps == 0
        Te = t_val - mean(P.t);                     % propagateParticleSmallDt.m:4-6
        dts = min(0.01*min(material.meanFreeTime(:)), Te);
        teff(it - 1) = Te;
        rec_begin('smalldt', 6 + 2*nbnd, N);
        P = propagateParticleSmallDt(material, geometry, P, t_val);
        draws{it - 1} = rec_end(ceil(Te/dts));
    else
        rec_begin('events', 6, N);
        P = propagateParticle(material, P, t_val);
        draws{it - 1} = rec_end(0);
    end
    E(:,:,it,:) = E(:,:,it,:) + observeTime(geometry, acoustics, P.x, P.p, P.dir, bins, ibins, vals, P.alive);
    states{it - 1} = state(P);
end
RTMC_REC.on = false;

mat = material_tables(material, geometry);
kind = 'smalldt';
if material.timeSteps ~= 0
    kind = 'events';
end
rec = struct('name', name, 'N', N, 'd', d, 'acoustics', double(acoustics), 'kind', kind, ...
             'geometry', geometry, 'source', source, 'observation', observation, 'mat', mat, ...
             'init_draws', init_draws, 'init_state', init_state, 'tval', tval, 'teff', teff, ...
             'E', E);
rec.draws = draws;
rec.states = states;
save(outfile, '-struct', 'rec', '-v7');
end

function rec_begin(kind, nslots, N)
global RTMC_REC
RTMC_REC.call = struct('kind', kind, 'nslots', nslots, 'N', N, 'nstep', 0, 'steps', {{}});
end

function buf = rec_end(kmin)
% the call's draws as a steps x slots x N array (NaN: not drawn); at least
% kmin steps (a sub-stepped call that ran out of particles breaks early)
global RTMC_REC
c = RTMC_REC.call;
K = max(numel(c.steps), kmin);
buf = nan(K, c.nslots, c.N);
for k = 1:numel(c.steps)
    if ~isempty(c.steps{k})
        buf(k, :, :) = reshape(c.steps{k}, 1, c.nslots, c.N);
    end
end
RTMC_REC.call = [];
end

function s = state(P)
s = struct('x', P.x, 'dir', P.dir, 'perp', P.perp, 'p', double(P.p), 'alive', double(P.alive), 't', P.t);
end

function m = material_tables(material, geometry)
m = struct('acoustics', double(material.acoustics), 'v', material.v, 'vp', material.vp, 'vs', material.vs, ...
           'meanFreeTime', material.meanFreeTime, 'P2P', material.P2P, 'S2S', material.S2S, ...
           'Q', material.Q, 'Frequency', material.Frequency, 'timeSteps', material.timeSteps);
if material.acoustics
    tabs = {material.invcdf};
else
    tabs = {material.invcdf{1,1}, material.invcdf{1,2}, material.invcdf{2,1}, material.invcdf{2,2}};
end
m.invcdf_x = cellfun(@(g) g.GridVectors{1}(:), tabs, 'UniformOutput', false);
m.invcdf_y = cellfun(@(g) g.Values(:), tabs, 'UniformOutput', false);
m.extrap = tabs{1}.ExtrapolationMethod;
m.zoe_deg = [];
m.zoe_Rpp = [];
m.zoe_Rsvsv = [];
if ~material.acoustics && isfield(geometry, 'bnd')
    [out, ~] = MaterialClass.Zoeppritz(material);
    m.zoe_deg = out.j1_deg(:);
    m.zoe_Rpp = out.E_Rpp(:);
    m.zoe_Rsvsv = out.E_Rsvsv(:);
end
end

function rtmc_stop(shadows)
global RTMC_REC
RTMC_REC = [];
rmpath(shadows);
end

function [geometry, source, material, observation] = rtmc_case(name, N)
switch name
    case {'readme3d', 'events3d'}
        geometry = struct('dimension', 3, 'frame', 'spherical');
        source = struct('numberParticles', N, 'type', 'point', 'position', [0 0 0], ...
                        'direction', 'uniform', 'lambda', 0.05);
        observation = struct('x', 0:0.05:6, 'y', [-pi pi], 'z', [-pi/2 pi/2], 'directions', [0 pi], ...
                             'time', 0:0.02:1);
        material = MaterialClass(geometry, 10, true, 1, [0.05 0.05], 0, 'exp', 0.1);
        if strcmp(name, 'readme3d')
            material.Q = 200;
        else
            material.timeSteps = 1;
        end
    case 'elastic2d'
        geometry = struct('dimension', 2, 'frame', 'cartesian');
        source = struct('numberParticles', N, 'type', 'point', 'position', [0 0], 'polarization', 'P', ...
                        'lambda', 0.002);
        observation = struct('x', -10:0.5:10, 'y', -10:0.5:10, 'time', 0:0.05:1);
        material = MaterialClass(geometry, 5, false, [6 3.46], [0.1 0.1 0.1], [0 0 0], 'gaussian', 0.5);
        material.Q = 500;
    case 'halfspace3d'
        geometry = struct('dimension', 3, 'frame', 'cylindrical', 'bnd', struct('dir', 3, 'val', 0));
        source = struct('numberParticles', N, 'type', 'point', 'position', [0 0 -0.5], 'lambda', 0.1);
        observation = struct('x', 0:0.1:6, 'directions', linspace(0, pi, 19), 'time', 0:0.05:1);
        material = MaterialClass(geometry, 20, false, [6 6/sqrt(3)], [0.05 0.05 0.05], [0 0 0], 'exp', 0.2);
        material.rho = 2000;
    otherwise
        error('rtmc_record:case', 'unknown case "%s"', name);
end
end

function m = rtmc_sites(refdir)
% Draw sites: file, source text, occurrence, call kind, slot, particle
% indices (evaluated in the caller, in draw order), selection of them,
% boundary number (1-based; 0 for none), step phase, reduction.
S = {
 'initializeParticle', 'r = abs(randn(N,1)*source.lambda);',        1, 'init', 0, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'r = invcdfsource(rand(N,1));',               1, 'init', 0, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'phi0 = acos(-1+2*rand(N,1));',               1, 'init', 1, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'phi = acos(-1+2*rand(N,1));',                1, 'init', 2, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'theta0 = 2*pi*rand(N,1);',                   1, 'init', 3, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'theta = 2*pi*rand(N,1);',                    1, 'init', 4, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'theta = 2*pi*rand(N,1);',                    2, 'init', 5, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'phi   = acos(rand(N,1));',                   1, 'init', 6, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'alpha = 2*pi*rand(N,1);',                    1, 'init', 7, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'c_long = randn(N,1) * source.lambda/2;',     1, 'init', 0, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'c_trans1 = (rand(N,1)-0.5) * L1;',           1, 'init', 1, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'c_trans2 = (rand(N,1)-0.5) * L2;',           1, 'init', 2, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'rho   = R * sqrt(rand(N,1));',               1, 'init', 1, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'alpha = 2*pi*rand(N,1);',                    2, 'init', 2, '(1:N)''', '', '0', 'init', ''
 'initializeParticle', 'theta = rand(N,1)*2*pi;',                    1, 'init', 4, '(1:N)''', '', '0', 'init', ''
 'propagateParticleSmallDt', 'kill(alive) = rand(nnz(alive),1) < pKillP;', 1, 'smalldt', 0, 'find(alive)', '', '0', 'pre', ''
 'propagateParticleSmallDt', 'kill(p) = rand(nnz(p),1) < pKillP;',   1, 'smalldt', 0, 'find(p)', '', '0', 'pre', ''
 'propagateParticleSmallDt', 'kill(s) = rand(nnz(s),1) < pKillS;',   1, 'smalldt', 0, 'find(s)', '', '0', 'pre', ''
 'propagateParticleSmallDt', 'rand_val = rand(length(glob_p), 1);',  1, 'smalldt', 6, 'glob_p', '', 'i2', 'pre', ''
 'propagateParticleSmallDt', 'rand_type = rand(length(glob_s), 1);', 1, 'smalldt', 6, 'glob_s', '', 'i2', 'pre', ''
 'propagateParticleSmallDt', 'rand_conv = rand(length(g_idx), 1);',  1, 'smalldt', 7, 'g_idx', '', 'i2', 'pre', ''
 'propagateParticleSmallDt', 'Nscat = poissrnd(lambda);',            1, 'smalldt', 1, '(1:P.N)''', '', '0', 'step', ''
 'propagateParticleSmallDt', 'phi = randi([0 1],Nscat,1,''logical'');', 1, 'smalldt', 2, 'find(scat)', '', '0', 'post', ''
 'propagateParticleSmallDt', 'phi = (2*pi)*rand(Nscat,1);',          1, 'smalldt', 2, 'find(scat)', '', '0', 'post', ''
 'propagateParticleSmallDt', 'theta(scat) = mat.invcdf(rand(Nscat,1));', 1, 'smalldt', 3, 'find(scat)', '', '0', 'post', ''
 'propagateParticleSmallDt', 'P.p( p & (rand(P.N,1)>mat.P2P) ) = false;', 1, 'smalldt', 4, '(1:P.N)''', '', '0', 'post', ''
 'propagateParticleSmallDt', 'P.p( s & (rand(P.N,1)>mat.S2S) ) = true;',  1, 'smalldt', 5, '(1:P.N)''', '', '0', 'post', ''
 'propagateParticleSmallDt', 'theta( scat ) = rand(Nscat,1);',       1, 'smalldt', 3, 'find(scat)', '', '0', 'post', ''
 'MaterialClass', 'isSV = rand(size(probSV)) < probSV;',             1, 'smalldt', 6, 'find(sIdx)', '', '1', 'pre', ''
 'MaterialClass', 'p2p  = rand(nP,1) <= Rpp_coeff;',                 1, 'smalldt', 6, 'find(pIdx)', '', '1', 'pre', ''
 'MaterialClass', 'sv2sv = rand(nSV,1) <= Rsvsv_coeff;',             1, 'smalldt', 7, 'find(sIdx)', 'isSV', '1', 'pre', ''
 'propagateParticle', 'Nj(p) = poissrnd(dt(p)./mat.meanFreeTime(1));', 1, 'events', 0, 'find(p)', '', '0', 'step', ''
 'propagateParticle', 'Nj(s) = poissrnd(dt(s)./mat.meanFreeTime(2));', 1, 'events', 0, 'find(s)', '', '0', 'post', ''
 'propagateParticle', 'tj(ind,i2) = rand(sum(ind),1).*dt(ind);',    1, 'events', 1, 'find(ind)', '', '0', 'post', 'minjump'
 'propagateParticle', 'phi = randi([0 1],Nscatter,1,''logical'');', 1, 'events', 2, 'find(scat)', '', '0', 'post', ''
 'propagateParticle', 'phi = (2*pi)*rand(Nscatter,1);',              1, 'events', 2, 'find(scat)', '', '0', 'post', ''
 'propagateParticle', 'theta = mat.invcdf(rand(Nscatter,1));',       1, 'events', 3, 'find(scat)', '', '0', 'post', ''
 'propagateParticle', 'P.p( p & (rand(P.N,1)>mat.P2P) ) = false;',   1, 'events', 4, '(1:P.N)''', '', '0', 'post', ''
 'propagateParticle', 'P.p( s & (rand(P.N,1)>mat.S2S) ) = true;',    1, 'events', 5, '(1:P.N)''', '', '0', 'post', ''
 'propagateParticle', 'theta = rand(P.N,1);',                        1, 'events', 3, '(1:P.N)''', '', '0', 'post', ''
};
files = struct('initializeParticle', fullfile(refdir, 'initializeParticle.m'), ...
               'propagateParticleSmallDt', fullfile(refdir, 'propagateParticleSmallDt.m'), ...
               'propagateParticle', fullfile(refdir, 'propagateParticle.m'), ...
               'MaterialClass', fullfile(refdir, '@MaterialClass', 'MaterialClass.m'));
m = containers.Map();
for i = 1:size(S, 1)
    lines = splitlines(fileread(files.(S{i,1})));
    hit = find(contains(strtrim(lines), S{i,2}));
    if numel(hit) < S{i,3}
        error('rtmc_record:site', 'draw site "%s" (occurrence %d) not found in %s', S{i,2}, S{i,3}, S{i,1});
    end
    line = hit(S{i,3});
    s = struct('text', S{i,2}, 'kind', S{i,4}, 'slot', S{i,5}, 'idx', S{i,6}, 'sub', S{i,7}, ...
               'bnd', S{i,8}, 'phase', S{i,9}, 'reduce', S{i,10});
    m(sprintf('%s:%d', S{i,1}, line)) = s;
end
end
