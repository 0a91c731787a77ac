function opts = checkBM3Dopts(opts)
% opts = checkBM3Dopts(opts) -- fills absent fields with the defaults of
% utilities/checkBM3Dopts.m.
opts = gbm3d_mex('checkBM3Dopts',opts);
