function [b_omw, OTF_bp_omw, abs_OTF_c, OTF_mask] = omw_backprojector_generation(psf, alpha, skewed, varargin)
% Drop-in for microscopeDataProcessing/decon/omw_backprojector_generation.m:
% same signature and outputs; the nearest-edge search runs on the GPU.
ip = inputParser;
ip.CaseSensitive = false;
ip.addRequired('psf');
ip.addRequired('alpha');
ip.addRequired('skewed');
ip.addParameter('OTFCumThresh', 0.9, @isnumeric);
ip.addParameter('OTFAreaThresh', 100, @isnumeric);
ip.addParameter('hannWinBounds', [0.8, 1], @(x) isnumeric(x) && numel(x) == 2);
ip.addParameter('minIntThrsh', 1e-3, @(x) isnumeric(x));
ip.parse(psf, alpha, skewed, varargin{:});
pr = ip.Results;
bo = struct('OTFCumThresh', pr.OTFCumThresh, 'OTFAreaThresh', pr.OTFAreaThresh, ...
    'hannWinBounds', double(pr.hannWinBounds), 'minIntThrsh', pr.minIntThrsh);
[b_omw, OTF_bp_omw, abs_OTF_c, OTF_mask] = pk_decon_mex('bp', single(psf), double(alpha), skewed, bo);
if ~any(OTF_mask(:) > 1), OTF_mask = OTF_mask > 0; else, OTF_mask = double(OTF_mask); end
end
