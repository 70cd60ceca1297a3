function c = minus(a, b)
% MINUS  Band-by-band difference a - b of two framelet expansions.
%   Replacement for matlab/classes/+DataTypes/@FrameletExpansion/minus.m
%   (same defects as plus.m).
    if a.dim ~= b.dim
        error('FrameletExpansion:dim', 'inconsistent dimensions: %d and %d', a.dim, b.dim);
    end
    c = DataTypes.FrameletExpansion(a.dim, a.frameletSystem, combineBands(a.frameletArray, b.frameletArray, @minus));
end
