function [PSF_bp, OTF_bp] = BackProjector(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes, verboseFlag)
% Drop-in for third_parties/WBDeconvolution/BackProjector.m: same signature,
% argument-count defaults and outputs (PSF_bp double, OTF_bp complex double),
% computed by libpkdecon.so on the host.  verboseFlag = 1 (log and figures)
% is served by the reference's own code, kept as BackProjector_reference.m.
if nargin < 2, bp_type = 'traditional'; end
if nargin < 3, alpha = 0.001; end
if nargin < 4, beta = 1; end
if nargin < 5, n = 10; end
if nargin < 6, resFlag = 1; iRes = [0, 0, 0]; end
if nargin < 7, iRes = [0, 0, 0]; end
if nargin < 8, verboseFlag = 0; end
if verboseFlag
    [PSF_bp, OTF_bp] = BackProjector_reference(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes, verboseFlag);
    return;
end
[PSF_bp, OTF_bp] = pk_decon_mex('wbbp', single(PSF_fp), bp_type, double(alpha), double(beta), double(n), ...
    double(resFlag), double(iRes));
end
