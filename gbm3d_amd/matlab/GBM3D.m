function [U,out] = GBM3D(I,sigma,opts)
% [U,out] = GBM3D(I,sigma,opts) -- drop-in for source/GBM3D.m of the
% reference, running on an AMD MI355X through libgbm3d (gbm3d_mex).
% I: m x n real image (double or single), sigma: scalar noise level on the
% image's scale, or a matrix sigmaPSD = sigma^2*PSD for coloured noise,
% opts: optional struct (see setBM3Dopts / checkBM3Dopts).
% U: denoised image (double), out.U1: the first-step estimate.
if nargin < 3
    opts.profile = 'default';
end
if size(sigma,1) > 1
    % coloured noise (source/GBM3D.m:33-35): the per-coefficient Wiener sigma
    % from the reference's getColoredWie where its makeGausPSF is on the
    % path, else [] (the library's getColoredWie, whose Gaussian stands in
    % for makeGausPSF)
    o = opts;
    if isfield(o,'profile'), o = setBM3Dopts(o.profile); end
    o = checkBM3Dopts(o);
    if exist('makeGausPSF','file') && exist('getColoredWie','file')
        sigmaWie = getColoredWie(sigma,o.blockSizeWie);
    else
        sigmaWie = [];
    end
    [U,out] = gbm3d_mex('GBM3D_colored',I,sigma,sigmaWie,opts);
else
    [U,out] = gbm3d_mex(I,sigma,opts);
end
