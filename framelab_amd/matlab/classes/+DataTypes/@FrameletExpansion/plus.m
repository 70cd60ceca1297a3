function c = plus(a, b)
% PLUS  Band-by-band sum of two framelet expansions of one framelet system.
%   Replacement for matlab/classes/+DataTypes/@FrameletExpansion/plus.m,
%   which reads properties the class does not define (nLevel,
%   frameletType) and calls a four-argument constructor.
    if a.dim ~= b.dim
        error('FrameletExpansion:dim', 'inconsistent dimensions: %d and %d', a.dim, b.dim);
    end
    c = DataTypes.FrameletExpansion(a.dim, a.frameletSystem, combineBands(a.frameletArray, b.frameletArray, @plus));
end
