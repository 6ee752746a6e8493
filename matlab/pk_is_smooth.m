function tf = pk_is_smooth(sz)
% true when every axis length factors into 2, 3 and 5 only (the sizes the
% GPU transforms take; see findGoodFactorNumber(n, 1, false) with max prime 5)
tf = true;
for n = sz
    f = factor(n);
    if n > 1 && max(f) > 5, tf = false; return; end
end
end
