"""Geometry modules of the projector API.

Mirrors deeprecon/deeprecon/torch/projector/projector.py: the same classes
(projector2d :55-216, projector3d :219-390, projectorsp :393-563), the same
constructor arguments, Options packing (:113, :286, :460; arc detectors get
det_size and binshift divided by source_detector, :110-112), shape asserts,
dispatch table (:8-53), FBP filter and cosine-weight generation and the FBP
composition (sino * weight -> conv with the filter -> weighted
backprojection -> * angular step, halved when redundant).

Two behavioural additions: the ops receive a host-side copy of ``options``
(refreshed whenever the parameter changes), so no op has to read geometry
back from the device, and the FBP filter stage (the reference's conv2d /
conv3d with 2D-1 taps and D-1 padding) runs as the equal product with the
D x D Toeplitz matrix of the taps, one library GEMM.  ``self.options`` stays
a non-trainable Parameter, as in the reference, and moves with the module.
"""
from typing import List, Tuple, Union

import torch
import torch.nn as nn
from torch import Tensor

from pridediff_amd.fbp import filter_sino

from . import function as F
from .filter import filter_generate

__functions__ = {
    "dis": {
        "flat": {
            "projection": F.ProjFanFlatDis2d,
            "projection_t": F.ProjTransFanFlatDis2d,
            "bprojection": F.BackProjFanFlatDis2d,
            "bprojection_t": F.BackProjTransFanFlatDis2d,
            "bprojection_w": F.BackProjWeightedFanFlatDis2d,
            "get_sys_matrix": F.GetSysMatrixFanFlat2d,
            "projection3d": F.ProjConeFlatDis3d,
            "projection3d_t": F.ProjTransConeFlatDis3d,
            "bprojection3d": F.BackProjConeFlatDis3d,
            "bprojection3d_t": F.BackProjTransConeFlatDis3d,
            "bprojection3d_w": F.BackProjWeightedConeFlatDis3d,
        },
        "arc": {
            "projection": F.ProjFanArcDis2d,
            "projection_t": F.ProjTransFanArcDis2d,
            "bprojection": F.BackProjFanArcDis2d,
            "bprojection_t": F.BackProjTransFanArcDis2d,
            "bprojection_w": F.BackProjWeightedFanArcDis2d,
            "get_sys_matrix": F.GetSysMatrixFanArc2d,
            "projection3d": F.ProjConeArcDis3d,
            "projection3d_t": F.ProjTransConeArcDis3d,
            "bprojection3d": F.BackProjConeArcDis3d,
            "bprojection3d_t": F.BackProjTransConeArcDis3d,
            "bprojection3d_w": F.BackProjWeightedConeArcDis3d,
            "projectionsp": F.ProjSpiralDis3d,
            "projectionsp_t": F.ProjTransSpiralDis3d,
            "bprojectionsp": F.BackProjSpiralDis3d,
            "bprojectionsp_t": F.BackProjTransSpiralDis3d,
            "bprojectionsp_w": F.BackProjWeightedSpiralDis3d,
        },
        "para": {
            "projection": F.ProjParaDis2d,
            "projection_t": F.ProjTransParaDis2d,
            "bprojection": F.BackProjParaDis2d,
            "bprojection_t": F.BackProjTransParaDis2d,
            "bprojection_w": F.BackProjParaDis2d,
            "get_sys_matrix": F.GetSysMatrixPara2d,
        },
    }
}


def _tuple(v, n, what):
    if not isinstance(v, (tuple, list)):
        v = (v,) * n
    assert len(v) == n, what
    return tuple(v)


class _ProjectorBase(nn.Module):
    """Shared plumbing: options parameter, host mirror, op dispatch."""

    def _init_common(self, options, dtype, filter_type, trainable):
        self.options = nn.parameter.Parameter(options, requires_grad=False)
        self._host_opts = None
        self._host_key = None
        self.filter = nn.parameter.Parameter(
            self.filter_generate(filter_type=self.filter_type, dtype=dtype), requires_grad=trainable)
        self.weight = nn.parameter.Parameter(self.weight_generate(dtype=dtype), requires_grad=trainable)

    def host_options(self) -> Tensor:
        """CPU float32 copy of ``self.options`` for the ops (no device reads
        unless the parameter changed since the last call)."""
        opt = self.options
        if opt.device.type == "cpu":
            return opt.detach()
        key = (opt.data_ptr(), opt._version)
        if self._host_key != key:
            self._host_opts = opt.detach().to("cpu", torch.float32)
            self._host_key = key
        return self._host_opts

    def _apply_op(self, name, x, angles, *extra):
        fn = __functions__[self.method][self.scan_type][name]
        return fn.apply(x, self.host_options(), angles, *extra)


class projector2d(_ProjectorBase):
    """2D projector (parallel, fan-flat or fan-arc) for CT reconstruction.

    Args match deeprecon's projector2d: image_size (int or (W, H)), num_det,
    pix_size (float or (dx, dy)), det_size, iso_source, source_detector,
    pixshift, binshift, scan_type ('flat' | 'arc' | 'para'), method ('dis'),
    filter_type ('ramp' | 'sheeplogan'), trainable, dtype.
    """

    def __init__(self, image_size: Union[int, Tuple[int], List[int]], num_det: int,
                 pix_size: Union[float, Tuple[float], List[float]], det_size: float,
                 iso_source: float, source_detector: float,
                 pixshift: Union[float, Tuple[float], List[float]] = 0.,
                 binshift: float = 0., scan_type: str = "flat", method: str = "dis",
                 filter_type: str = "ramp", trainable: bool = False, dtype=None) -> None:
        super().__init__()
        image_size = _tuple(image_size, 2, "Image resolution should be a int or two-tuple int")
        pix_size = _tuple(pix_size, 2, "Pixle size should be a float or two-tuple float")
        pixshift = _tuple(pixshift, 2, "Scanned object shift should be a float or two-tuple float")
        scan_type, method, filter_type = scan_type.lower(), method.lower(), filter_type.lower()
        assert scan_type in {"flat", "arc", "para"}, f"Scanning mode {scan_type} not supported"
        assert method in {"dis"}, f"Radon transform method {method} not supported"
        if scan_type == "arc":
            det_size = det_size / source_detector
            binshift = binshift / source_detector
        options = torch.tensor([*image_size, num_det, *pix_size, det_size, iso_source,
                                source_detector, *pixshift, binshift], dtype=dtype)
        self.image_size = tuple(image_size)
        self.num_det = num_det
        self.det_size = det_size
        self.binshift = binshift
        self.iso_source = iso_source
        self.source_detector = source_detector
        self.scan_type = scan_type
        self.method = method
        self.filter_type = filter_type
        self._init_common(options, dtype, filter_type, trainable)

    def _check_image(self, image):
        b, c, h, w = image.shape
        assert (w, h) == self.image_size, "The size of input image is inconsistent with geometry"

    def _check_sino(self, sino, angles):
        b, c, h, w = sino.shape
        v, = angles.shape
        assert (h, w) == (v, self.num_det), "The size of input sinogram is inconsistent with geometry"

    def projection(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._apply_op("projection", image, angles)

    def backprojection(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("bprojection", sino, angles)

    def backprojection_w(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("bprojection_w", sino, angles)

    def projection_t(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("projection_t", sino, angles)

    def backprojection_t(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._apply_op("bprojection_t", image, angles)

    def forward(self, image: Tensor, angles: Tensor) -> Tensor:
        return self.projection(image, angles)

    def filter_generate(self, filter_type, dtype=None):
        det = self.det_size * self.iso_source / self.source_detector if self.scan_type == "flat" \
            else self.det_size
        return filter_generate(self.num_det, det, filter_type, dtype)[None, None, None, :]

    def weight_generate(self, dtype=None):
        u = torch.arange(-self.num_det / 2 + 0.5, self.num_det / 2, 1, dtype=dtype)
        u = u * self.det_size + self.binshift
        if self.scan_type == "flat":
            u = u * self.iso_source / self.source_detector
            return self.iso_source / torch.sqrt(u ** 2 + self.iso_source ** 2)
        if self.scan_type == "arc":
            return torch.cos(u)
        return torch.ones_like(u)

    def filtered_backprojection(self, sino: Tensor, angles: Tensor,
                                filter: Union[Tensor, None] = None,
                                filter_type: Union[str, None] = None,
                                redundant: bool = True) -> Tensor:
        if filter is not None:
            taps = filter
        elif filter_type is None or filter_type.lower() == self.filter_type:
            taps = self.filter
        else:
            taps = self.filter_generate(filter_type=filter_type.lower(), dtype=sino.dtype).to(sino.device)
        filtered = filter_sino(sino, self.weight, taps)
        recon = self.backprojection_w(filtered, angles)
        recon = recon * (angles[-1] - angles[0]).abs() / (angles.size(0) - 1)
        return recon / 2 if redundant else recon

    def get_sys_matrix(self, angles: Tensor) -> Tensor:
        return __functions__[self.method][self.scan_type]["get_sys_matrix"](self.options, angles)


class _Projector3dBase(_ProjectorBase):
    """Filter and weight generation shared by the 3D projectors
    (projector.py:350-371, :523-544)."""

    def filter_generate(self, filter_type, dtype=None):
        det = self.det_size[0] * self.iso_source / self.source_detector if self.scan_type == "flat" \
            else self.det_size[0]
        return filter_generate(self.num_det[0], det, filter_type, dtype)[None, None, None, None, :]

    def weight_generate(self, dtype=None):
        col_size, row_size = self.det_size
        col_n, row_n = self.num_det
        col_s, row_s = self.binshift
        R, S = self.iso_source, self.source_detector
        wr = (torch.arange(-row_n / 2 + 0.5, row_n / 2, 1, dtype=dtype) * row_size + row_s) * R / S
        wc = torch.arange(-col_n / 2 + 0.5, col_n / 2, 1, dtype=dtype) * col_size + col_s
        if self.scan_type == "flat":
            wc = wc * R / S
        elif self.scan_type == "arc":
            wc = torch.tan(wc) * R
        wr = wr[:, None].repeat(1, col_n)
        wc = wc[None, :].repeat(row_n, 1)
        return R / torch.sqrt(wr ** 2 + wc ** 2 + R ** 2)


class projector3d(_Projector3dBase):
    """3D circular cone-beam projector ('flat' or 'arc' detector).

    image_size (W, H, Dz), num_det (cols, rows), pix_size (dx, dy, dz),
    det_size (col, row), pixshift (sx, sy, sz), binshift (col, row).
    """

    def __init__(self, image_size, num_det, pix_size, det_size, iso_source: float,
                 source_detector: float, pixshift=0., binshift=0., scan_type: str = "flat",
                 method: str = "dis", filter_type: str = "ramp", trainable: bool = False,
                 dtype=None) -> None:
        super().__init__()
        image_size = _tuple(image_size, 3, "Image resolution should be a int or three-tuple int")
        num_det = _tuple(num_det, 2, "Number of detector elements should be a int or two-tuple int")
        pix_size = _tuple(pix_size, 3, "Pixle size should be a float or three-tuple float")
        det_size = _tuple(det_size, 2, "The size of detector element should be a float or two-tuple float")
        pixshift = _tuple(pixshift, 3, "Scanned object shift should be a float or three-tuple float")
        binshift = _tuple(binshift, 2, "Detector shift should be a float or two-tuple float")
        scan_type, method, filter_type = scan_type.lower(), method.lower(), filter_type.lower()
        assert scan_type in {"flat", "arc"}, f"Scanning mode {scan_type} not supported"
        assert method in {"dis"}, f"Radon transform method {method} not supported"
        if scan_type == "arc":
            det_size = (det_size[0] / source_detector, det_size[1])
            binshift = (binshift[0] / source_detector, binshift[1])
        options = torch.tensor([*image_size, *num_det, *pix_size, *det_size, iso_source,
                                source_detector, *pixshift, *binshift], dtype=dtype)
        self.image_size = tuple(image_size)
        self.num_det = tuple(num_det)
        self.det_size = det_size
        self.binshift = binshift
        self.iso_source = iso_source
        self.source_detector = source_detector
        self.scan_type = scan_type
        self.method = method
        self.filter_type = filter_type
        self._init_common(options, dtype, filter_type, trainable)

    def _check_image(self, image):
        b, c, d, h, w = image.shape
        assert (w, h, d) == self.image_size, "The size of input image is inconsistent with geometry"

    def _check_sino(self, sino, angles):
        b, c, d, h, w = sino.shape
        v, = angles.shape
        assert (d, w, h) == (v, *self.num_det), "The size of input sinogram is inconsistent with geometry"

    def projection(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._apply_op("projection3d", image, angles)

    def backprojection(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("bprojection3d", sino, angles)

    def backprojection_w(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("bprojection3d_w", sino, angles)

    def projection_t(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino(sino, angles)
        return self._apply_op("projection3d_t", sino, angles)

    def backprojection_t(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._apply_op("bprojection3d_t", image, angles)

    def forward(self, image: Tensor, angles: Tensor) -> Tensor:
        return self.projection(image, angles)

    def filtered_backprojection(self, sino: Tensor, angles: Tensor,
                                filter: Union[Tensor, None] = None,
                                filter_type: Union[str, None] = None,
                                redundant: bool = True) -> Tensor:
        if filter is not None:
            taps = filter
        elif filter_type is None or filter_type.lower() == self.filter_type:
            taps = self.filter
        else:
            taps = self.filter_generate(filter_type=filter_type.lower(), dtype=sino.dtype).to(sino.device)
        filtered = filter_sino(sino, self.weight, taps)
        recon = self.backprojection_w(filtered, angles)
        recon = recon * (angles[-1] - angles[0]).abs() / (angles.size(0) - 1)
        return recon / 2 if redundant else recon


class projectorsp(_Projector3dBase):
    """Helical (spiral) cone-beam projector, curved ('arc') detector.

    image_size (W, H, Dz), num_det (cols, rows), pix_size (dx, dy, dz),
    det_size (col, row), pixshift (sx, sy), binshift (col, row);
    ``source_pos`` is [4, V]: axial position, radius shift, angle shift,
    axial focal-spot shift.
    """

    def __init__(self, image_size, num_det, pix_size, det_size, iso_source: float,
                 source_detector: float, pixshift=0., binshift=0., scan_type: str = "arc",
                 method: str = "dis", filter_type: str = "ramp", trainable: bool = False,
                 dtype=None) -> None:
        super().__init__()
        image_size = _tuple(image_size, 3, "Image resolution should be a int or three-tuple int")
        num_det = _tuple(num_det, 2, "Number of detector elements should be a int or two-tuple int")
        pix_size = _tuple(pix_size, 3, "Pixle size should be a float or three-tuple float")
        det_size = _tuple(det_size, 2, "The size of detector element should be a float or two-tuple float")
        pixshift = _tuple(pixshift, 2, "Scanned object shift should be a float or two-tuple float")
        binshift = _tuple(binshift, 2, "Detector shift should be a float or two-tuple float")
        scan_type, method, filter_type = scan_type.lower(), method.lower(), filter_type.lower()
        assert scan_type in {"arc"}, f"Scanning mode {scan_type} not supported"
        assert method in {"dis"}, f"Radon transform method {method} not supported"
        det_size = (det_size[0] / source_detector, det_size[1])
        binshift = (binshift[0] / source_detector, binshift[1])
        options = torch.tensor([*image_size, *num_det, *pix_size, *det_size, iso_source,
                                source_detector, *pixshift, *binshift], dtype=dtype)
        self.image_size = tuple(image_size)
        self.num_det = tuple(num_det)
        self.det_size = det_size
        self.binshift = binshift
        self.iso_source = iso_source
        self.source_detector = source_detector
        self.scan_type = scan_type
        self.method = method
        self.filter_type = filter_type
        self._init_common(options, dtype, filter_type, trainable)

    def _check(self, image_or_sino, angles, source_pos, sino):
        b, c, d, h, w = image_or_sino.shape
        v, = angles.shape
        m, n = source_pos.shape
        if sino:
            assert (d, w, h) == (v, *self.num_det), "The size of input sinogram is inconsistent with geometry"
        else:
            assert (w, h, d) == self.image_size, "The size of input image is inconsistent with geometry"
        assert (m, n) == (4, v), "The size of angles and source position is not consistent"

    def projection(self, image: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        self._check(image, angles, source_pos, False)
        return self._apply_op("projectionsp", image, angles, source_pos)

    def backprojection(self, sino: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        self._check(sino, angles, source_pos, True)
        return self._apply_op("bprojectionsp", sino, angles, source_pos)

    def backprojection_w(self, sino: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        self._check(sino, angles, source_pos, True)
        return self._apply_op("bprojectionsp_w", sino, angles, source_pos)

    def projection_t(self, sino: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        self._check(sino, angles, source_pos, True)
        return self._apply_op("projectionsp_t", sino, angles, source_pos)

    def backprojection_t(self, image: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        self._check(image, angles, source_pos, False)
        return self._apply_op("bprojectionsp_t", image, angles, source_pos)

    def forward(self, image: Tensor, angles: Tensor, source_pos: Tensor) -> Tensor:
        return self.projection(image, angles, source_pos)

    def filtered_backprojection(self, sino: Tensor, angles: Tensor, source_pos: Tensor,
                                filter_type: Union[str, None] = None) -> Tensor:
        if filter_type is None or filter_type.lower() == self.filter_type:
            taps = self.filter
        else:
            taps = self.filter_generate(filter_type=filter_type.lower(), dtype=sino.dtype).to(sino.device)
        filtered = filter_sino(sino, self.weight, taps)
        recon = self.backprojection_w(filtered, angles, source_pos)
        return recon * abs(angles[1] - angles[0]) / 2
