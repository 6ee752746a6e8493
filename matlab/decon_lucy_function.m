function [J, err_mat, k] = decon_lucy_function(I, PSF, NUMIT, varargin)
% Drop-in for microscopeDataProcessing/decon/decon_lucy_function.m (same
% signature, name-value options and outputs) served by the MI355X engine
% through pk_decon_mex.  Cases the GPU engine does not serve (useGPU=false,
% double input, debug=true, non-2/3/5-smooth sizes) go to the reference's
% own implementation, kept on the path as decon_lucy_function_reference.m.
ip = inputParser;
ip.CaseSensitive = false;
ip.addRequired('I', @isnumeric);
ip.addRequired('PSF', @isnumeric);
ip.addRequired('NUMIT', @isnumeric);
ip.addParameter('Background', [], @isnumeric);
ip.addParameter('useGPU', true, @islogical);
ip.addParameter('save16bit', true, @islogical);
ip.addParameter('dampFactor', 1, @isnumeric);
ip.addParameter('scaleFactor', 1, @isnumeric);
ip.addParameter('deconOffset', 0, @isnumeric);
ip.addParameter('EdgeErosion', 0, @isnumeric);
ip.addParameter('bbox', [], @isnumeric);
ip.addParameter('debug', false, @islogical);
ip.addParameter('debug_folder', './debug/', @ischar);
ip.addParameter('saveStep', 1, @isnumeric);
ip.parse(I, PSF, NUMIT, varargin{:});
pr = ip.Results;
if ~pr.useGPU || isa(I, 'double') || pr.debug || ~pk_is_smooth(size(I))
    [J, err_mat, k] = decon_lucy_function_reference(I, PSF, NUMIT, varargin{:});
    return;
end
opts = struct('Background', pr.Background, 'save16bit', pr.save16bit, 'dampFactor', pr.dampFactor, ...
    'scaleFactor', pr.scaleFactor, 'deconOffset', pr.deconOffset, 'EdgeErosion', pr.EdgeErosion, 'bbox', pr.bbox);
if ~isa(I, 'uint16'), I = single(I); end
[J, err_mat] = pk_decon_mex('rl', I, single(PSF), NUMIT, opts);
k = NUMIT;
end
