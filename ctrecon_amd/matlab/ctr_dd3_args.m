function [geo, dx, dz, gmask] = ctr_dd3_args(cfg, mask)
% CTR_DD3_ARGS  Geometry arguments of the DD3_GPU gateway for every view of cfg.
%   [geo, dx, dz, gmask] = ctr_dd3_args(cfg, mask) returns geo, a cell array of the 17
%   geometry arguments that DD3_GPU('Proj' | 'Back' | 'OSSART' | 'CG', ...) takes after
%   the action name (x0 y0 z0 DNU DNV xds yds zds xoff yoff zoff angles zshifts PN XN YN
%   ZN), the voxel sizes dx, dz and the mask bytes gmask, computed the way
%   iterativeRecon/dd3.m:86-186 computes them for its GPU path: source at (0, sid, 0)
%   (sid * para_scale in parallel mode), arc detector (or flat, cfg.recon.flat_det),
%   helical z shifts, view angles + pi/2 and the x / y offsets swapped
%   (dd3.m:247-250).  The revo-gap geometry and the column upsampling of dd3.m are not
%   supported.
acq = cfg.acq;
rec = cfg.recon;
x0 = single(0);
if isfield(rec, 'para_mode') && rec.para_mode
    if isfield(rec, 'para_scale')
        y0 = single(acq.sid * rec.para_scale);
    else
        y0 = single(acq.sid * 1e10);
    end
else
    y0 = single(acq.sid);
end
z0 = single(0);
nviews = acq.num_data_views;
ang = (0:(nviews - 1)) * 2 * pi / single(acq.views_per_rotation) + acq.first_view_angle / 180 * pi;
DNU = acq.total_col_count;
DNV = acq.row_count;
cols = 1:single(DNU);
if isfield(rec, 'flat_det') && rec.flat_det
    xds = (cols - (rec.col_center + 1)) * acq.col_width;
    yds = (acq.sid - acq.sdd) * ones(size(xds), 'single');
else
    alphas = (cols - (rec.col_center + 1)) * acq.col_width / acq.sdd;
    xds = sin(alphas) * acq.sdd;
    yds = acq.sid - acq.sdd * cos(alphas);
end
zds = ((1:single(DNV)) - single(DNV + 1) * 0.5) * acq.col_height;
zsh = (ang - acq.first_view_angle / 180 * pi) / 2 / pi * single(acq.helical_pitch) ...
      * acq.col_height_at_iso * single(DNV) + acq.first_view_zposition;
dx = single(rec.dfov_mm / single(max(rec.recon_pixels_x, rec.recon_pixels_y)));
dz = single(rec.recon_slice_spacing);
% no target centre (dd3.m:168-173): cen_x = cen_y = 0; the GPU path swaps them
% (xoffset = -cen_y, yoffset = cen_x, dd3.m:249-250)
cen_x = 0;
cen_y = 0;
zoff = single(rec.recon_center_z);
if nargin < 2 || isempty(mask)
    gmask = int8(ones(rec.recon_pixels_x, rec.recon_pixels_y, 2));
else
    gmask = int8([mask, mask']);
end
geo = {x0, y0, z0, int32(DNU), int32(DNV), single(xds), single(yds), single(zds), ...
       single(-cen_y), single(cen_x), zoff, single(ang + pi / 2), single(zsh), int32(nviews), ...
       int32(rec.recon_pixels_x), int32(rec.recon_pixels_y), int32(rec.recon_pixels_z)};
end
