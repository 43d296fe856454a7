function [reconImg, resid] = conjgrad(initImg, prj, cfg, mask, iterNum)
% CONJGRAD  Conjugate gradients on Back(Proj(x)) = Back(prj) with the whole loop on the GPU.
%   reconImg = conjgrad(initImg, prj, cfg, mask, iterNum) keeps the signature and the
%   iteration of iterativeRecon/conjgrad.m:1-35 (A = Back(Proj(.)), b = Back(prj),
%   r0 = b - A x0, p0 = r0, alpha = r'r / p'Ap, beta = r1'r1 / r0'r0, no stopping
%   rule) with the reference back projector as Back.  Unlike the reference it makes one
%   call, DD3_GPU('CG', ...): the vectors stay on the GPU, the dot products and step
%   scalars are double precision, and nothing is drawn.
%   [reconImg, resid] = conjgrad(...) also returns ||r_k|| of every iteration.
%
%   initImg  single [recon_pixels_z, recon_pixels_x, recon_pixels_y]
%   prj      single [row_count, total_col_count, num_data_views]
%   mask     [recon_pixels_x, recon_pixels_y] (nonzero inside the field of view)
[geo, dx, dz, gmask] = ctr_dd3_args(cfg, mask);
img = permute(single(initImg), [1 3 2]);  % the GPU's axis order (dd3.m:248)
[x, resid] = DD3_GPU('CG', geo{:}, img, single(prj), dx, dz, gmask, 0, int32(iterNum), ...
                     int32(0));
reconImg = permute(x, [1 3 2]);  % back to [z, x, y] (dd3.m:317)
end
