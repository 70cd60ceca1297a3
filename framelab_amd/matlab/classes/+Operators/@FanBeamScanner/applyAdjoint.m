function Aty = applyAdjoint(this, data)
% APPLYADJOINT  Fan-beam backprojection A^T y of a DataTypes.CTData object.
%   Replacement for matlab/classes/+Operators/@FanBeamScanner/applyAdjoint.m,
%   whose check reads data.para, a field CTData does not have.  The scan
%   parameters come from the scanner that produced the data (data.scanner),
%   as ConeBeamScanner/applyAdjoint.m:8-14 takes them.  The gateway
%   Atx_fan_mf runs the MI355X library (libframelab_hip) when a GPU is
%   present, the CPU kernels otherwise.
    src = data.scanner.para;
    if isempty(fieldnames(this.para))
        this.para = src;
    elseif ~structcmp(this.para, src)
        error('FanBeamScanner:para', 'Data scan parameters and scanner parameters do not agree!');
    end
    Y = data.dataArray;
    if gpuDeviceCount > 0
        this.para.GPU = uint32(1);
        this.checkInputs(Y);
        x = Atx_fan_mf(Y, this.para);
    else
        this.para.GPU = uint32(0);
        this.checkInputs(Y);
        x = Atx_fan_mf_cpu(Y, this.para);
    end
    if this.verbose
        fprintf('fan-beam adjoint, %d x %d image, GPU=%d\n', this.para.nx, this.para.ny, this.para.GPU);
    end
    Aty = DataTypes.ObjectData(2, reshape(x, [this.para.nx this.para.ny]), data.L);
end
