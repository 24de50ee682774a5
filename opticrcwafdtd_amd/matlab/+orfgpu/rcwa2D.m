function out = rcwa2D(devs, lambda0, theta, phi, pol, opts)
%RCWA2D orf.rcwa2D on an MI355X (gfx950) through liborf_rcwa.so.
%   out = orfgpu.rcwa2D(devs, lambda0, theta, phi, pol, Name=Value) has the
%   signature, options and output struct of orf.rcwa2D
%   (toolbox/+orf/rcwa2D.m:2-35); every solve runs on the GPU through the
%   orf_rcwa_mex gateway.  UseGPU / GPUEig are accepted and ignored.
%   Extensions that default to the reference's behaviour: a vector lambda0
%   (wavelength as a trailing output dimension), a cell array of pols
%   (polarisation as a further trailing dimension) and Precision.
arguments
    devs
    lambda0 (1,:) double {mustBePositive}
    theta (1,1) double
    phi (1,1) double
    pol
    opts.Nx (1,1) double {mustBeInteger, mustBeNonnegative} = 5
    opts.Ny double = []
    opts.Nsample (1,:) double = 0
    opts.Factorization (1,1) string {mustBeMember(opts.Factorization, ["li" "laurent"])} = "li"
    opts.GapEps (1,1) double = NaN
    opts.UseGPU (1,1) logical = true
    opts.GPUEig (1,1) logical = true
    opts.BatchSize (1,1) double = 0
    opts.Symmetry (1,1) string {mustBeMember(opts.Symmetry, ["auto" "none"])} = "auto"
    opts.Progress = []
    opts.Precision (1,1) string {mustBeMember(opts.Precision, ["fp64" "mixed"])} = "fp64"
end
if isstruct(devs), devs = {devs}; end
B = numel(devs);
period = devs{1}.period;
Ny = opts.Ny; if isempty(Ny), Ny = opts.Nx; end
if isinf(period(2)), Ny = 0; end
L = numel(lambda0);

% polarisations: "s" / "p" / [aTE aTM], or a cell array of them
if iscell(pol), pols = pol; else, pols = {pol}; end
polKind = zeros(1, numel(pols), 'int32');
jones = zeros(4, numel(pols));
for i = 1:numel(pols)
    q = pols{i};
    if (isstring(q) || ischar(q)) && any(strcmp(q, ["s" "p"]))
        polKind(i) = int32(strcmp(q, "p"));
    elseif isnumeric(q) && numel(q) == 2
        polKind(i) = 2;
        jones(:, i) = [real(q(1)); imag(q(1)); real(q(2)); imag(q(2))];
    else
        error('rcwa:pol', 'pol must be "s", "p" or [aTE aTM].');
    end
end

% materials: eps = n^2 per (material, wavelength), deduplicated by value
matEps = complex(zeros(0, L));
    function k = mid(mat)
        e = reshape(mat.nk(lambda0), 1, []).^2;
        hit = find(all(matEps == e, 2), 1);
        if isempty(hit)
            matEps(end + 1, :) = e;
            k = int32(size(matEps, 1) - 1);
        else
            k = int32(hit - 1);
        end
    end
inc = mid(devs{1}.incident);
sub = mid(devs{1}.substrate);
nL = numel(devs{1}.layers);
thickness = zeros(nL, B);
background = zeros(nL, B, 'int32');
shapeStart = zeros(nL, B, 'int32');
shapeCount = zeros(nL, B, 'int32');
shapes = zeros(0, 8);
vertices = zeros(2, 0);
for b = 1:B
    d = devs{b};
    if ~isequal(d.period, period) || numel(d.layers) ~= nL
        error('orf:invalidArgument', 'structures must share the period and the layer count');
    end
    for j = 1:nL
        lay = d.layers{j};
        thickness(j, b) = lay.thickness;
        background(j, b) = mid(lay.background);
        shapeStart(j, b) = size(shapes, 1);
        shapeCount(j, b) = numel(lay.shapes);
        for s = 1:numel(lay.shapes)
            sh = lay.shapes{s};
            row = [0, double(mid(sh.mat)), sh.center(1), sh.center(2), 0, 0, 0, 0];
            switch sh.type
                case "circle", row(1) = 0; row(5) = sh.radius;
                case "rect", row(1) = 1; row(5:6) = sh.size;
                case "polygon"
                    row(1) = 2; row(7) = size(vertices, 2); row(8) = size(sh.vertices, 1);
                    vertices = [vertices, sh.vertices.'];  %#ok<AGROW>
                otherwise, error('orf:invalidArgument', 'unsupported shape "%s"', sh.type);
            end
            shapes(end + 1, :) = row;  %#ok<AGROW>
        end
    end
end

% sampling: the automatic rule of this library (edge-aligned, NsMin = max(256, 8(4N+1)))
Ns = opts.Nsample;
if isscalar(Ns) && Ns == 0
    Ns = [orf.alignedSampleCount(devs, period, 1, max(256, 8 * (4 * opts.Nx + 1))), 1];
    if isfinite(period(2)), Ns(2) = orf.alignedSampleCount(devs, period, 2, max(256, 8 * (4 * Ny + 1))); end
end

p = struct('Lx', period(1), 'Ly', period(2), 'Nx', opts.Nx, 'Ny', Ny, 'Nsx', Ns(1), 'Nsy', Ns(end), ...
    'lambda0', lambda0, 'theta', theta, 'phi', phi, 'polKind', polKind, 'jones', jones, ...
    'matEps', complex(matEps), 'incMaterial', inc, 'subMaterial', sub, 'nStruct', B, 'nLayers', nL, ...
    'thickness', thickness, 'background', background, 'shapeStart', shapeStart, ...
    'shapeCount', shapeCount, 'shapes', shapes, 'vertices', vertices, ...
    'factorization', double(opts.Factorization == "laurent"), 'symmetry', double(opts.Symmetry == "none"), ...
    'gapEps', opts.GapEps, 'precision', double(opts.Precision == "mixed"), 'batchSize', opts.BatchSize);
raw = orf_rcwa_mex(p, opts.Progress);

% MATLAB shapes of the reference: R, T M x B (x L x P); r, t 3 x M x B (x L x P)
M = numel(raw.m);
P = numel(pols);
out.m = double(raw.m);
out.n = double(raw.n);
out.R = squeezeTrailing(reshape(raw.R, M, B, L, P));
out.T = squeezeTrailing(reshape(raw.T, M, B, L, P));
out.Rtot = squeezeTrailing(reshape(raw.Rtot, 1, B, L, P));
out.Ttot = squeezeTrailing(reshape(raw.Ttot, 1, B, L, P));
out.r = squeezeTrailing(reshape(raw.r, 3, M, B, L, P));
out.t = squeezeTrailing(reshape(raw.t, 3, M, B, L, P));
end

function a = squeezeTrailing(a)
% drop trailing singleton wavelength / polarisation dimensions only
sz = size(a);
while numel(sz) > 2 && sz(end) == 1, sz(end) = []; end
a = reshape(a, [sz, ones(1, max(0, 2 - numel(sz)))]);
end
