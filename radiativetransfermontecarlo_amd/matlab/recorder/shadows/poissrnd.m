function r = poissrnd(varargin)
%POISSRND  Recording shadow of the Statistics Toolbox poissrnd, on the path only
%while rtmc_record runs: calls the toolbox function (handle taken by
%rtmc_record before this folder went on the path) and stores the counts of
%the reference's draw sites.  The toolbox's own rand/randn calls come from
%poissrnd's file, which is no draw site, so they are not stored.
global RTMC_REC
if isempty(RTMC_REC)
    error('rtmc_record:inactive', 'this poissrnd shadow works only while rtmc_record runs');
end
r = RTMC_REC.poissrnd(varargin{:});
s = rtmc_rec_site();
if ~isempty(s)
    idx = evalin('caller', s.idx);
    if ~isempty(s.sub)
        idx = idx(evalin('caller', s.sub));
    end
    rtmc_rec_put(s, idx, r, evalin('caller', s.bnd));
end
end
