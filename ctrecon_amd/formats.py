"""On-disk scan and image formats of the reference GUI, and its cfg adapter.

``ProjectionData`` follows ProjectionData.m:1-176 and ``ReconImage`` ReconImage.m:1-118:
the same fields, defaults, derived geometry and binary files (float64, little endian,
MATLAB column-major element order):

    ProjectionData file: SourceToCenterDistance, SourceToDetectorDistance, DetResoU,
        DetResoV, DetPixelSizeU, DetPixelSizeV, DetCntIdxU, DetCntIdxV, IsArcDetector,
        StartSourceAngle (degrees), StartSourceZPos, Pitch, TotalProjectionNumber,
        ProjectionNumberPerTurn, projection (its stored dimensions), StorageOrder (3)
        (ProjectionData.m:113-131 / 137-172)
    ReconImage file: resoX, resoY, resoZ, pixelSizeX, pixelSizeY, pixelSizeZ, ctrIdxX,
        ctrIdxY, ctrIdxZ, imgCenterX, imgCenterY, imgCenterZ, image, storageOrder (3)
        (ReconImage.m:60-118)

StorageOrder gives the axis order of the stored array in terms of (U, V, P) for
projections and (X, Y, Z) for images: [2, 1, 3] means the array is [V, U, P].

``to_cfg`` is wrapperFromProjectionAndReconImageToCfg.m:1-28, including its quirks:
the scan's Pitch (mm per turn in ProjectionData.zshiftCallBack) is passed as dd3's
unitless ``helical_pitch``, and the image centre is not passed (recon_center_z = 0).
``reconstruct`` is the data path of ossartGUI.m / conjugateGradientGUI.m without the
GUI: cfg from the files, the FOV mask of ossartGUI.m:5-15, the permutes of
ossartGUI.m:18-22 / 122, and the GPU OS-SART or CG.
"""

from __future__ import annotations

import math

import numpy as np

__all__ = ["ProjectionData", "ReconImage", "to_cfg", "reconstruct"]

_ORDERS = [(1, 2, 3), (2, 1, 3), (3, 1, 2), (1, 3, 2), (2, 3, 1), (3, 2, 1)]


def _canonical(arr, order):
    """Array stored with axis order ``order`` -> canonical (1, 2, 3) order."""
    order = tuple(int(o) for o in order)
    if order not in _ORDERS:
        raise ValueError(f"unsupported storage order {order}")
    # stored axis i holds canonical axis order[i] - 1
    return np.transpose(arr, np.argsort(np.array(order) - 1))


def _stored(arr, order):
    """Canonical (1, 2, 3) array -> array with axis order ``order``."""
    return np.transpose(arr, np.array(order, dtype=int) - 1)


def _read_f64(path):
    return np.fromfile(path, dtype="<f8")


class ProjectionData:
    """ProjectionData.m: helical scan geometry and projections."""

    def __init__(self):
        # defaults of the reference constructor (ProjectionData.m:37-71)
        self.SourceToCenterDistance = 538.0
        self.SourceToDetectorDistance = 946.0
        self.DetResoU = 888
        self.DetResoV = 64
        self.DetPixelSizeU = 1.0239
        self.DetPixelSizeV = 1.0963
        self.DetCntIdxU = 443.75  # zero-based
        self.DetCntIdxV = 31.5    # zero-based
        self.IsArcDetector = True
        self.StartSourceAngle = 0.0  # degrees
        self.StartSourceZPos = 0.0
        self.Pitch = 0.0             # mm per turn
        self.TotalProjectionNumber = 984
        self.ProjectionNumberPerTurn = 984
        self.StorageOrder = (1, 2, 3)
        self.projection = np.zeros((self.DetResoU, self.DetResoV, self.TotalProjectionNumber))
        self._views = None
        self._zshifts = None

    # ---- derived geometry (ProjectionData.m:352-379) ----
    @property
    def source(self):
        return 0.0, float(self.SourceToCenterDistance), 0.0

    def view_angles(self):
        """Radians: k 2 pi / ProjectionNumberPerTurn + StartSourceAngle (viewAnglesCallBack);
        replaced by set_view_angles, as getSubProjectionSet does."""
        if self._views is not None:
            return self._views
        k = np.arange(int(self.TotalProjectionNumber), dtype=np.float64)
        return k * 2 * math.pi / self.ProjectionNumberPerTurn + self.StartSourceAngle / 180 * math.pi

    def zshifts(self):
        """(angle - StartSourceAngle) / (2 pi) * Pitch + StartSourceZPos (zshiftCallBack)."""
        if self._zshifts is not None:
            return self._zshifts
        a = self.view_angles()
        return (a - self.StartSourceAngle / 180 * math.pi) / (2 * math.pi) * self.Pitch + \
            self.StartSourceZPos

    def set_view_angles(self, a):
        self._views = np.asarray(a, np.float64)

    def set_zshifts(self, z):
        self._zshifts = np.asarray(z, np.float64)

    def xyz_ds(self):
        """Detector cell centres (xyzDsCallBack): arc (or flat) columns, rows along z."""
        u = np.arange(1, int(self.DetResoU) + 1, dtype=np.float64) - (self.DetCntIdxU + 1)
        if self.IsArcDetector:
            a = u * self.DetPixelSizeU / self.SourceToDetectorDistance
            x = np.sin(a) * self.SourceToDetectorDistance
            y = self.SourceToCenterDistance - self.SourceToDetectorDistance * np.cos(a)
        else:
            x = u * self.DetPixelSizeU
            y = np.full_like(x, self.SourceToCenterDistance - self.SourceToDetectorDistance)
        z = (np.arange(1, int(self.DetResoV) + 1) - self.DetCntIdxV - 1) * self.DetPixelSizeV
        return x, y, z

    # ---- projections ----
    def projection_uvp(self):
        """The projections as a [U, V, P] array whatever the storage order."""
        return _canonical(np.asarray(self.projection), self.StorageOrder)

    def set_projection(self, prj, order=(1, 2, 3)):
        self.projection = np.asarray(prj)
        self.StorageOrder = tuple(int(o) for o in order)

    def dd3_sinogram(self):
        """[V, U, P] float32, the layout dd3 / os_sart take (ossartGUI.m:21-22)."""
        return np.ascontiguousarray(np.transpose(self.projection_uvp(), (1, 0, 2)),
                                    dtype=np.float32)

    def sub_projection_set(self, osNum: int, osIdx: int) -> "ProjectionData":
        """getSubProjectionSet (ProjectionData.m:74-93): views osIdx, osIdx + osNum, ...
        (osIdx one-based, as in the reference)."""
        res = ProjectionData()
        for f in ("SourceToCenterDistance", "SourceToDetectorDistance", "DetResoU", "DetResoV",
                  "DetPixelSizeU", "DetPixelSizeV", "DetCntIdxU", "DetCntIdxV",
                  "IsArcDetector", "Pitch"):
            setattr(res, f, getattr(self, f))
        views, zs = self.view_angles(), self.zshifts()
        res.StartSourceAngle = float(views[osIdx - 1])
        res.StartSourceZPos = float(zs[osIdx - 1])
        idx = np.arange(osIdx - 1, int(self.TotalProjectionNumber), osNum)
        res.TotalProjectionNumber = int(idx.size)
        res.ProjectionNumberPerTurn = int(self.ProjectionNumberPerTurn // osNum)
        res.set_projection(self.projection_uvp()[:, :, idx])
        res.set_view_angles(views[idx])
        res.set_zshifts(zs[idx])
        return res

    # ---- files ----
    def write(self, path):
        head = [self.SourceToCenterDistance, self.SourceToDetectorDistance, self.DetResoU,
                self.DetResoV, self.DetPixelSizeU, self.DetPixelSizeV, self.DetCntIdxU,
                self.DetCntIdxV, float(bool(self.IsArcDetector)), self.StartSourceAngle,
                self.StartSourceZPos, self.Pitch, self.TotalProjectionNumber,
                self.ProjectionNumberPerTurn]
        body = np.asarray(self.projection, dtype=np.float64).ravel(order="F")
        data = np.concatenate([np.asarray(head, np.float64), body,
                               np.asarray(self.StorageOrder, np.float64)])
        data.astype("<f8").tofile(path)

    @classmethod
    def read(cls, path) -> "ProjectionData":
        d = _read_f64(path)
        o = cls()
        (o.SourceToCenterDistance, o.SourceToDetectorDistance, U, V, o.DetPixelSizeU,
         o.DetPixelSizeV, o.DetCntIdxU, o.DetCntIdxV, arc, o.StartSourceAngle,
         o.StartSourceZPos, o.Pitch, P, PT) = (float(x) for x in d[:14])
        o.DetResoU, o.DetResoV, o.TotalProjectionNumber = int(U), int(V), int(P)
        o.ProjectionNumberPerTurn = int(PT) if float(PT).is_integer() else float(PT)
        o.IsArcDetector = bool(arc)
        n = o.DetResoU * o.DetResoV * o.TotalProjectionNumber
        if d.size != 14 + n + 3:
            raise ValueError(f"{path}: {d.size} doubles, expected {14 + n + 3}")
        order = tuple(int(x) for x in d[14 + n:])
        dims = {1: o.DetResoU, 2: o.DetResoV, 3: o.TotalProjectionNumber}
        shape = tuple(dims[a] for a in order)
        o.set_projection(d[14:14 + n].reshape(shape, order="F"), order)
        return o


class ReconImage:
    """ReconImage.m: image grid and image."""

    def __init__(self):
        # defaults of the reference constructor (ReconImage.m:31-52)
        self.resoX, self.resoY, self.resoZ = 512, 512, 64
        self.pixelSizeX = self.pixelSizeY = self.pixelSizeZ = 0.9766
        self.imgCenterX = self.imgCenterY = self.imgCenterZ = 0.0
        self.ctrIdxX = (self.resoX - 1) / 2
        self.ctrIdxY = (self.resoY - 1) / 2
        self.ctrIdxZ = (self.resoZ - 1) / 2
        self.storageOrder = (1, 2, 3)
        self.image = np.zeros((self.resoX, self.resoY, self.resoZ))

    # the centre index / centre position pair is kept consistent as in the setters of
    # ReconImage.m:121-211: ctrIdx = (reso - 1) / 2 - imgCenter / pixelSize
    def set_img_center(self, x=None, y=None, z=None):
        for ax, v in (("X", x), ("Y", y), ("Z", z)):
            if v is not None:
                setattr(self, "imgCenter" + ax, float(v))
                setattr(self, "ctrIdx" + ax, (getattr(self, "reso" + ax) - 1) / 2 -
                        float(v) / getattr(self, "pixelSize" + ax))

    def set_ctr_idx(self, x=None, y=None, z=None):
        for ax, v in (("X", x), ("Y", y), ("Z", z)):
            if v is not None:
                setattr(self, "ctrIdx" + ax, float(v))
                setattr(self, "imgCenter" + ax, ((getattr(self, "reso" + ax) - 1) / 2 -
                                                 float(v)) * getattr(self, "pixelSize" + ax))

    def image_xyz(self):
        return _canonical(np.asarray(self.image), self.storageOrder)

    def set_image(self, img, order=(1, 2, 3)):
        self.image = np.asarray(img)
        self.storageOrder = tuple(int(o) for o in order)

    def dd3_image(self):
        """[Z, X, Y] float32, the layout dd3 / os_sart take (ossartGUI.m:18-19)."""
        return np.ascontiguousarray(np.transpose(self.image_xyz(), (2, 0, 1)), dtype=np.float32)

    def set_from_dd3(self, img_zxy):
        """Inverse of dd3_image (ossartGUI.m:122-123: permute [2 3 1])."""
        self.set_image(np.transpose(np.asarray(img_zxy, np.float64), (1, 2, 0)))

    def write(self, path):
        head = [self.resoX, self.resoY, self.resoZ, self.pixelSizeX, self.pixelSizeY,
                self.pixelSizeZ, self.ctrIdxX, self.ctrIdxY, self.ctrIdxZ, self.imgCenterX,
                self.imgCenterY, self.imgCenterZ]
        body = np.asarray(self.image, dtype=np.float64).ravel(order="F")
        data = np.concatenate([np.asarray(head, np.float64), body,
                               np.asarray(self.storageOrder, np.float64)])
        data.astype("<f8").tofile(path)

    @classmethod
    def read(cls, path) -> "ReconImage":
        d = _read_f64(path)
        o = cls()
        (X, Y, Z, o.pixelSizeX, o.pixelSizeY, o.pixelSizeZ, o.ctrIdxX, o.ctrIdxY, o.ctrIdxZ,
         o.imgCenterX, o.imgCenterY, o.imgCenterZ) = (float(x) for x in d[:12])
        o.resoX, o.resoY, o.resoZ = int(X), int(Y), int(Z)
        n = o.resoX * o.resoY * o.resoZ
        if d.size != 12 + n + 3:
            raise ValueError(f"{path}: {d.size} doubles, expected {12 + n + 3}")
        order = tuple(int(x) for x in d[12 + n:])
        dims = {1: o.resoX, 2: o.resoY, 3: o.resoZ}
        o.set_image(d[12:12 + n].reshape(tuple(dims[a] for a in order), order="F"), order)
        return o


def to_cfg(pd: ProjectionData, img: ReconImage) -> dict:
    """wrapperFromProjectionAndReconImageToCfg.m:1-28."""
    sid, sdd = float(pd.SourceToCenterDistance), float(pd.SourceToDetectorDistance)
    return {
        "acq": {
            "sid": sid,
            "sdd": sdd,
            "views_per_rotation": pd.ProjectionNumberPerTurn,
            "num_data_views": int(pd.TotalProjectionNumber),
            "row_count": int(pd.DetResoV),
            "total_col_count": int(pd.DetResoU),
            "col_width": float(pd.DetPixelSizeU),
            "col_height": float(pd.DetPixelSizeV),
            "rotation_direction": 1,
            "first_view_angle": float(pd.StartSourceAngle),
            "first_view_zposition": float(pd.StartSourceZPos),
            "helical_pitch": float(pd.Pitch),
            "col_height_at_iso": sid / sdd * float(pd.DetPixelSizeV),
        },
        "recon": {
            "col_center": float(pd.DetCntIdxU),
            "row_center": float(pd.DetCntIdxV),
            "recon_pixels_z": int(img.resoZ),
            "recon_pixels_y": int(img.resoY),
            "recon_pixels_x": int(img.resoX),
            "dfov_mm": float(img.pixelSizeX) * int(img.resoX),
            "recon_slice_spacing": float(img.pixelSizeZ),
            "recon_center_z": 0.0,
            "flat_det": not bool(pd.IsArcDetector),
        },
    }


def gui_mask(img: ReconImage) -> np.ndarray:
    """The FOV mask of ossartGUI.m:5-15 ([resoX, resoY], int8, one-based ii / jj)."""
    xx, yy = int(img.resoX), int(img.resoY)
    ii = np.arange(1, xx + 1, dtype=np.float64)[:, None]
    jj = np.arange(1, yy + 1, dtype=np.float64)[None, :]
    m = ((ii - 0.5 - xx / 2) / (xx / 2)) ** 2 + ((jj - 0.5 - yy / 2) / (yy / 2)) ** 2 < 0.99
    return m.astype(np.int8)


def reconstruct(pd: ProjectionData, img: ReconImage, method="ossart", osNum=10, iterNum=10,
                bp_kind="ref") -> ReconImage:
    """ossartGUI / conjugateGradientGUI without the GUI: the image in ``img`` is the
    initial value and receives the result (the whole loop runs on the GPU)."""
    from . import dd3 as D

    cfg = to_cfg(pd, img)
    mask = gui_mask(img)
    x0 = img.dd3_image()
    prj = pd.dd3_sinogram()
    if method == "ossart":
        out = D.os_sart(x0, prj, cfg, mask, osNum, iterNum, bp=bp_kind)
    elif method == "cg":
        out = D.conjgrad(x0, prj, cfg, mask, iterNum, bp=bp_kind)
    else:
        raise ValueError("method must be 'ossart' or 'cg'")
    img.set_from_dd3(out)
    return img
