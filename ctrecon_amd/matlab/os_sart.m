function [reconImg, resid] = os_sart(initImg, prj, cfg, mask, osNum, iterNum)
% OS_SART  Ordered-subset SART with the whole loop on the GPU.
%   reconImg = os_sart(initImg, prj, cfg, mask, osNum, iterNum) keeps the signature and
%   the iteration of iterativeRecon/os_sart.m:1-44: subset s (s = 0 .. osNum-1) holds
%   the views s+1 : osNum : end; RowSum_s = A_s 1 and ColSum_s = B_s 1 are computed once,
%   and each sub-iteration is x = x + B_s((prj_s - A_s x) ./ RowSum_s) ./ ColSum_s, the
%   divisions taken where the sums are nonzero (A: the branchless distance-driven
%   forward projector, B: its reference back projector).  Unlike the reference it makes
%   one call, DD3_GPU('OSSART', ...): the geometry tables, the sums and the iterate
%   stay on the GPU for all iterNum sweeps, and nothing is drawn.
%   [reconImg, resid] = os_sart(...) also returns the relative data residual
%   ||prj - A x|| / ||prj|| accumulated over each sweep.
%
%   initImg  single [recon_pixels_z, recon_pixels_x, recon_pixels_y]
%   prj      single [row_count, total_col_count, num_data_views]
%   mask     [recon_pixels_x, recon_pixels_y] (nonzero inside the field of view)
[geo, dx, dz, gmask] = ctr_dd3_args(cfg, mask);
img = permute(single(initImg), [1 3 2]);  % the GPU's axis order (dd3.m:248)
[x, resid] = DD3_GPU('OSSART', geo{:}, img, single(prj), dx, dz, gmask, 0, ...
                     int32(osNum), int32(iterNum), int32(0));
reconImg = permute(x, [1 3 2]);  % back to [z, x, y] (dd3.m:317)
end
