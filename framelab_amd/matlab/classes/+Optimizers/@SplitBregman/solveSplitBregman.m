function u = solveSplitBregman(obj)
% SOLVESPLITBREGMAN  Split Bregman for min_u 0.5|A u - f|^2 + lam |W u|_1
%   (low-pass bands unpenalised), the class form of
%   matlab/classes/+Optimizers/@SplitBregman/solveSplitBregman.m:16-41:
%     alpha = W u0, v = alpha, b = A^T f + mu W^T (alpha - v);
%     repeat: u = CGLS on (A^T A + mu I) u = b (cgIter steps, warm start);
%             alpha = shrink(v + W u, lam/mu); v = v + (W u - alpha);
%             b = A^T f + mu W^T (alpha - v).
%   Differs from the reference only in keeping softThreshold's result
%   (alpha = alpha.softThreshold(...)).  With the MI355X library the same
%   recurrence runs device-resident as fl_split_bregman.
    alpha = obj.W(obj.u0);
    v = alpha;
    Atf = obj.At(obj.f);
    b = Atf + obj.mu * obj.Wt(alpha - v);
    u = obj.u0;
    for k = 1:obj.globalIter
        cg = Optimizers.CGLS(obj.A, obj.At, b, u, obj.mu, obj.cgIter, obj.cgTol);
        u = cg.solveCGLS();
        Wu = obj.W(u);
        alpha = v + Wu;
        alpha = alpha.softThreshold(obj.lam / obj.mu);
        v = v + (Wu - alpha);
        b = Atf + obj.mu * obj.Wt(alpha - v);
    end
end
