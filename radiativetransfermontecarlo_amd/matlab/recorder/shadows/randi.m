function r = randi(varargin)
%RANDI  Recording shadow of the builtin randi, on the path only while
%rtmc_record runs: returns the builtin's values and,
%when the caller is one of the reference's draw sites, stores them in
%librtmc's replay layout.  Calls with a RandStream dispatch to its method and
%are not seen here.
r = builtin('randi', varargin{:});
s = rtmc_rec_site();
if ~isempty(s)
    idx = evalin('caller', s.idx);
    if ~isempty(s.sub)
        idx = idx(evalin('caller', s.sub));
    end
    rtmc_rec_put(s, idx, r, evalin('caller', s.bnd));
end
end
