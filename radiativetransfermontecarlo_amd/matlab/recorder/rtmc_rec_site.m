function s = rtmc_rec_site()
%RTMC_REC_SITE  The draw site (struct) of the function that called the
%random-number function calling this one, or [] when no recording runs or
%that caller is none of the reference's draw sites (sites are keyed by file
%name and line, located by rtmc_record from their source text).
global RTMC_REC
s = [];
if isempty(RTMC_REC) || ~RTMC_REC.on || isempty(RTMC_REC.call)
    return
end
st = dbstack('-completenames');   % this function, the shadow, its caller
if numel(st) < 3
    return
end
[~, f] = fileparts(st(3).file);
key = sprintf('%s:%d', f, st(3).line);
if isKey(RTMC_REC.sites, key)
    s = RTMC_REC.sites(key);
end
end
