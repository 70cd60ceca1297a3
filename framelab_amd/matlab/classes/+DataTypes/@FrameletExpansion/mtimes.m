function c = mtimes(a, b)
% MTIMES  Scalar times framelet expansion (either order).
%   Replacement for matlab/classes/+DataTypes/@FrameletExpansion/mtimes.m
%   (same defects as plus.m).
    if isnumeric(a) && isscalar(a)
        c = DataTypes.FrameletExpansion(b.dim, b.frameletSystem, combineBands(b.frameletArray, a, @times));
    elseif isnumeric(b) && isscalar(b)
        c = DataTypes.FrameletExpansion(a.dim, a.frameletSystem, combineBands(a.frameletArray, b, @times));
    else
        error('FrameletExpansion:mtimes', 'only a scalar times a framelet expansion is defined');
    end
end
