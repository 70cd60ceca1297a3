function obj = softThreshold(obj, tau)
% SOFTTHRESHOLD  sign(a) .* max(|a| - tau, 0) on every band of every level
%   except the low-pass band {1,1} / {1,1,1}, which is kept.
%   Replacement for matlab/classes/+DataTypes/@FrameletExpansion/
%   softThreshold.m: FrameletExpansion is a value class, so the method must
%   return the shrunk object (the reference's returns nothing and its
%   caller, solveSplitBregman.m:33, keeps the unshrunk coefficients).  Call
%   as  alpha = alpha.softThreshold(tau).
    A = obj.frameletArray;
    for L = 1:numel(A)
        bands = A{L};
        for b = 2:numel(bands)          % linear index 1 is the low-pass band
            a = bands{b};
            bands{b} = sign(a) .* max(abs(a) - tau, 0);
        end
        A{L} = bands;
    end
    obj = DataTypes.FrameletExpansion(obj.dim, obj.frameletSystem, A);
end
