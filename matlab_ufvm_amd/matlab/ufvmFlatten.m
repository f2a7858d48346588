function v = ufvmFlatten(c)
% ufvmFlatten  The rows of a cell array (theMesh.cconn, coefficients.anb)
% concatenated in row order into one column, whatever each row's orientation.
if isempty(c)
    v = zeros(0, 1);
    return
end
parts = cellfun(@(x) reshape(x, [], 1), c(:), 'UniformOutput', false);
v = double(vertcat(parts{:}));
end
