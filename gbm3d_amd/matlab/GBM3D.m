function [U,out] = GBM3D(I,sigma,opts)
% [U,out] = GBM3D(I,sigma,opts) -- drop-in for source/GBM3D.m of the
% reference, running on an AMD MI355X through libgbm3d (gbm3d_mex).
% I: m x n real image (double or single), sigma: scalar noise level on the
% image's scale, opts: optional struct (see setBM3Dopts / checkBM3Dopts).
% U: denoised image (double), out.U1: the first-step estimate.
if nargin < 3
    [U,out] = gbm3d_mex(I,sigma);
else
    [U,out] = gbm3d_mex(I,sigma,opts);
end
