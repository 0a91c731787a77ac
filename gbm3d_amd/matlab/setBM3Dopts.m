function opts = setBM3Dopts(profile)
% opts = setBM3Dopts(profile) -- profiles 'default', 'fast', 'superFast',
% 'accuracy' (unknown names give 'default'), as utilities/setBM3Dopts.m.
if nargin < 1, profile = 'default'; end
opts = gbm3d_mex('setBM3Dopts',profile);
