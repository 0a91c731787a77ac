function [U,out] = GBM3D_distributed(I,sigma,opts)
% [U,out] = GBM3D_distributed(I,sigma,opts) -- drop-in for
% source/GBM3D_distributed.m of the reference: overlapping tiles of
% 256 + blockSize denoised independently and stitched with ramp windows,
% on an AMD MI355X through libgbm3d (gbm3d_mex).
% U: denoised image (double), out.U2: the first-step estimate.
if nargin < 3
    opts.profile = 'default';
end
[U,out] = gbm3d_mex('GBM3D_distributed',I,sigma,opts);
