function compile_framelab_amd(framelabRoot, bins)
%COMPILE_FRAMELAB_AMD Build FrameLab's mex gateways against libframelab_hip.
%   compile_framelab_amd(framelabRoot) compiles the four gateways
%   Ax_fan_mf, Atx_fan_mf, Ax_cone_mf and Atx_cone_mf from
%   <framelabRoot>/matlab/mex/src and writes the mex files to
%   <framelabRoot>/matlab/mex, where the Operators classes call them.
%   The gateways and their CPU kernels are compiled unchanged; the GPU
%   entry points they declare extern "C" (Ax_fan_mf.h, Atx_fan_mf.h,
%   Ax_cone_mf.h, Atx_cone_mf.h) resolve to libframelab_hip.so, which
%   exports them with the same parameter lists, in place of the .cu
%   objects and -lcudart of matlab/compile_framelab.m:39-93.
%
%   compile_framelab_amd(framelabRoot, bins) builds a subset, e.g.
%   compile_framelab_amd(root, {'Ax_cone_mf','Atx_cone_mf'}).
%
%   Build libframelab_hip.so first (make -C framelab_amd/csrc).  The mex
%   files find it, and the ROCm runtime, through an rpath.

if nargin < 2 || isempty(bins)
    bins = {'Ax_fan_mf', 'Atx_fan_mf', 'Ax_cone_mf', 'Atx_cone_mf'};
end
here = fileparts(mfilename('fullpath'));          % <repo>/framelab_amd/matlab
libDir = fileparts(here);                          % <repo>/framelab_amd
incDir = fullfile(fileparts(libDir), 'include');   % <repo>/include
rocmLib = getenv('ROCM_PATH');
if isempty(rocmLib), rocmLib = '/opt/rocm'; end
rocmLib = fullfile(rocmLib, 'lib');
if ~isfile(fullfile(libDir, 'libframelab_hip.so'))
    error('compile_framelab_amd:nolib', 'libframelab_hip.so not found in %s (build it first)', libDir);
end
src = fullfile(framelabRoot, 'matlab', 'mex', 'src');
out = fullfile(framelabRoot, 'matlab', 'mex');

% each gateway with the CPU kernels it still dispatches to when para.GPU = 0
sources.Ax_fan_mf = {'Ax_fan_mf.cpp', 'Ax_fan_mf_cpu_siddon.cpp', 'Ax_fan_mf_cpu_new.cpp', ...
                     'Ax_fan_mf_cpu_new_fb.cpp', 'find_area.cpp', 'sort_alpha.cpp'};
sources.Atx_fan_mf = {'Atx_fan_mf.cpp', 'Atx_fan_mf_cpu_new.cpp', 'Atx_fan_mf_cpu_new_fb.cpp', ...
                      'find_l.cpp', 'find_area.cpp'};
sources.Ax_cone_mf = {'Ax_cone_mf.cpp', 'Ax_cone_mf_cpu_siddon.cpp', 'Ax_cone_mf_cpu_new.cpp', ...
                      'sort_alpha.cpp'};
sources.Atx_cone_mf = {'Atx_cone_mf.cpp', 'Atx_cone_mf_cpu_new.cpp', 'find_l_3d.cpp'};

ldflags = sprintf('LDFLAGS=$LDFLAGS -Wl,-rpath,%s -Wl,-rpath,%s', libDir, rocmLib);
for i = 1:numel(bins)
    name = bins{i};
    if ~isfield(sources, name)
        error('compile_framelab_amd:bin', 'unknown gateway %s', name);
    end
    files = cellfun(@(f) fullfile(src, f), sources.(name), 'UniformOutput', false);
    fprintf('compile_framelab_amd: %s\n', name);
    mex('-outdir', out, ['-I' src], ['-I' incDir], files{:}, ['-L' libDir], '-lframelab_hip', ldflags);
end
end
