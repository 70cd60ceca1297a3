function C = combineBands(A, B, op)
% COMBINEBANDS  Applies op band by band to two framelet arrays of the same
%   shape ({level}{i,j} or {level}{i,j,k}); B may be a scalar, applied to
%   every band.  Helper of plus, minus and mtimes (private to the class's methods).
    C = A;
    for L = 1:numel(A)
        if isnumeric(B)
            C{L} = cellfun(@(a) op(a, B), A{L}, 'UniformOutput', false);
        else
            if ~isequal(size(A{L}), size(B{L}))
                error('FrameletExpansion:shape', 'framelet arrays differ in band layout at level %d', L);
            end
            C{L} = cellfun(op, A{L}, B{L}, 'UniformOutput', false);
        end
    end
end
