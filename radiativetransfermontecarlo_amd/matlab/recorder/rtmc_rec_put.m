function rtmc_rec_put(s, idx, v, b)
%RTMC_REC_PUT  Stores the values v of draw site s, one per particle idx (in
%the order the site draws them), in the running call's replay buffer at
%slot s.slot (boundary b, 1-based: slots 6 + 2(b-1) and 7 + 2(b-1)) of its
%step: every draw of initializeParticle is step 1; a sub-step or loop
%iteration starts at its 'step' site, and 'pre' sites (removals, boundaries)
%draw before it within the same sub-step.
global RTMC_REC
if ~strcmp(RTMC_REC.call.kind, s.kind)
    return
end
nstep = RTMC_REC.call.nstep;
nslots = RTMC_REC.call.nslots;
N = RTMC_REC.call.N;
switch s.phase
    case 'init'
        k = 1;
    case 'step'
        k = nstep + 1;
        RTMC_REC.call.nstep = k;
    case 'pre'
        k = nstep + 1;
    otherwise
        k = nstep;
end
slot = s.slot;
if slot >= 6
    slot = slot + 2*(b - 1);
end
idx = double(idx(:));
v = double(v(:));
if numel(idx) ~= numel(v)
    error('rtmc_record:size', 'draw site "%s": %d values for %d particles', s.text, numel(v), numel(idx));
end
if k > numel(RTMC_REC.call.steps) || isempty(RTMC_REC.call.steps{k})
    RTMC_REC.call.steps{k} = nan(nslots, N);
end
if strcmp(s.reduce, 'minjump')
    % timeNextJump (propagateParticle.m:86-96): idx are positions among the
    % particles with a positive event count this iteration, and slot 1 keeps
    % the minimum of each particle's uniforms (tj = min(u) dt, dt > 0)
    cnt = RTMC_REC.call.steps{k}(1, :);
    cnt(isnan(cnt)) = 0;
    sc = find(cnt > 0);
    idx = sc(idx);
    old = RTMC_REC.call.steps{k}(slot + 1, idx);
    old(isnan(old)) = inf;
    v = min(old(:), v);
end
RTMC_REC.call.steps{k}(slot + 1, idx) = v;
end
