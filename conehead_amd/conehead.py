"""Conehead.calculate on the GPU engine (drop-in for conehead/conehead.py:520).

The reference's flow -- density interpolation, hit test, effective depth,
off-axis distance, fluence, softening and horn, TERMA, polyenergetic kernel,
dose (conehead/conehead.py:522-773) -- runs as one fused TERMA launch and
the collapsed-cone convolution (five FFT passes, or one gather launch),
with every intermediate kept in HBM.  The
dose is returned and stored in self.dose_grid.dose (the reference plots it
and discards it).

Settings keys are the reference's (example.py:27-58).  Optional keys:
  thetas, phis, kernel -- the cone set handed to the convolution (default:
      12 azimuths and all 48 polar keys of PolyenergeticKernel, the cone set
      of the reference's convolve_c call, conehead/conehead.py:780-794);
  semantics -- "convolve_c" (default), "cuda_dose" (y-polar cones and the
      index arithmetic of conehead/conehead.py:385-488; `kernel` is then a
      [n_phi, >=5996] cumulative table and `phis` its angles; without them
      the reference's live call: the EDKnrc KernelPoly of
      settings['energy_weights'], its 48 polar angles and 12 float32
      azimuths, conehead/conehead.py:686-738) or "physics"
      (the cuda_dose cones aligned with the beam and the kernel looked up at
      the radiological distance through the densities; without `kernel` the
      EDKnrc polyenergetic kernel of settings['energy_weights'] merged into
      settings.get('n_phi', 12) polar groups);
  convolution -- "fft" (default for the convolve_c and cuda_dose
      semantics: the cone stencil summed into one kernel and applied with
      zero-padded FFTs, ch_convolve_fft) or "gather" (the direct stencil
      sum, ch_convolve); physics mode always gathers, since its kernel
      distance depends on the densities along each path;
  density_xy_swap -- True (default) reproduces the reference's density
      resampling, whose meshgrid(indexing='xy') evaluation at the grid nodes
      swaps the x and y axes (conehead/conehead.py:562-567); False uses the
      densities as given.
"""
import numpy as np

from . import engine as _engine
from .cones import ConeSet, beam_frame, default_thetas, kernel_phis
from .batch import spectral_ok
from .device import DeviceArray, Stream
from .dosegrid import DoseGrid
from .kernel import EdkSet, merge_polar
from .kernel2 import PolyenergeticKernel


def dose_grid_densities(phantom, swap_xy=True):
    """Densities on the dose grid (the phantom's own grid).

    The reference interpolates the phantom linearly at the nodes
    linspace(origin, origin + n*s, n) of the same axes it interpolates on,
    so every value is a node value; its meshgrid uses indexing='xy', which
    lays the result out as [y][x][z]."""
    d = np.asarray(phantom.densities, dtype=np.float32)
    if not swap_xy:
        return np.ascontiguousarray(d)
    if d.shape[0] != d.shape[1]:
        raise ValueError(
            "the reference's xy-swapped density resampling is only defined "
            "for phantoms with size[0] == size[1]; pass "
            "settings['density_xy_swap'] = False")
    return np.ascontiguousarray(d.transpose(1, 0, 2))


class Conehead:

    def __init__(self, engine=None):
        self._engine = engine
        self._stream = None
        self.dose_grid = None

    @property
    def stream(self):
        """The instance's stream, kept across calls so its TERMA flag, FFT
        work grids and gather workspace are allocated once."""
        if self._stream is None:
            self._stream = Stream()
        return self._stream

    @property
    def engine(self):
        if self._engine is None:
            from .convolve_c import default_engine
            self._engine = default_engine()
        return self._engine

    def calculate(self, source, block, phantom, settings):
        """TERMA and dose of one beam; returns the FP32 dose grid.

        The dose grid is the phantom's own (conehead/conehead.py:536) unless
        settings['dose_grid'] = (size, origin, spacing) names another one;
        the densities are then resampled on the GPU (ch_resample_linear)
        with the reference's interpolation and layout.  A CtPhantom's
        densities are made on the GPU from its CT numbers
        (ch_ct_to_density) and resampled there."""
        eng = self.engine
        swap = settings.get("density_xy_swap", True)
        grid = settings.get("dose_grid")
        s = self.stream
        on_device = hasattr(phantom, "device_densities")  # CtPhantom
        if grid is None and not on_device:
            self.dose_grid = DoseGrid(phantom.size, phantom.origin,
                                      phantom.spacing)
            dens = dose_grid_densities(phantom, swap)
            d_rho = DeviceArray.from_host(dens, stream=s)
        else:
            # resampling the phantom's own grid onto itself gives every
            # node's value, i.e. dose_grid_densities bit for bit
            from .density import resample
            self.dose_grid = DoseGrid(*grid) if grid is not None else \
                DoseGrid(phantom.size, phantom.origin, phantom.spacing)
            if on_device:
                d_src = phantom.device_densities(stream=s)
            else:
                d_src = DeviceArray.from_host(phantom.densities, stream=s)
            d_rho = resample(d_src, tuple(phantom.size), phantom.origin,
                             phantom.spacing, self.dose_grid.size,
                             self.dose_grid.origin, self.dose_grid.spacing,
                             swap_xy=swap, stream=s)
            dens = d_rho.download(stream=s)
        shape = tuple(int(v) for v in self.dose_grid.size)
        origin = np.asarray(self.dose_grid.origin, dtype=np.float64)
        spacing = np.asarray(self.dose_grid.spacing, dtype=np.float64)
        if dens.shape != shape:
            raise ValueError("densities %s do not match size %s" %
                             (dens.shape, shape))

        semantics = settings.get("semantics", "convolve_c")
        kernel = settings.get("kernel")
        physics = semantics == "physics"
        if semantics == "convolve_c":
            kernel = kernel if kernel is not None else PolyenergeticKernel()
            phis = settings.get("phis")
            phis = kernel_phis(kernel) if phis is None else phis
        elif physics and kernel is None:
            # EDKnrc polyenergetic kernel of the beam's spectrum
            # (conehead/conehead.py:686-704), merged into polar groups
            poly = EdkSet.load().poly(settings["energy_weights"])
            phis, kernel = merge_polar(poly.angles, poly.kernel_cum,
                                       int(settings.get("n_phi", 12)))
        elif semantics == "cuda_dose" and kernel is None:
            # the reference's live dose call (conehead/conehead.py:686-738):
            # KernelPoly of settings['energy_weights'] (48 x 5996), its 48
            # float32 polar angles and 12 float32 azimuths 0, 30, ..., 330
            if settings.get("phis") is not None:
                raise ValueError("cuda_dose semantics: settings['phis'] "
                                 "needs settings['kernel'] too")
            poly = EdkSet.load().poly(settings["energy_weights"])
            kernel = poly.kernel_cum
            phis = np.asarray(poly.angles)
        else:
            if kernel is None or settings.get("phis") is None:
                raise ValueError("%s semantics needs settings['kernel'] "
                                 "(cumulative table) and settings['phis']"
                                 % semantics)
            phis = settings["phis"]
        thetas = settings.get("thetas")
        if thetas is None:
            thetas = (np.linspace(0, 360 - 360 / 12, 12, dtype=np.float32)
                      if semantics == "cuda_dose" else default_thetas(12))
        cones = ConeSet(thetas, phis, kernel, semantics=semantics,
                        frame=beam_frame(source) if physics else None)

        method = settings.get("convolution", "fft")
        if method not in ("fft", "gather"):
            raise ValueError("settings['convolution'] must be 'fft' or "
                             "'gather'")
        # the FFT form takes axes of up to 512 voxels; beyond, the gather
        use_fft = method == "fft" and not physics and spectral_ok(shape)
        d_terma = DeviceArray(shape, np.float32)
        # the support box only serves the gathers, which clip to it
        d_box = None if use_fft else DeviceArray((6,), np.int32)
        d_raster = eng.aperture(block, s)
        beam = _engine.beam_struct(source, settings, d_raster, d_rho,
                                   d_terma, d_box)
        _engine.terma([beam], shape, origin, spacing, stream=s)
        plan = eng.plan(shape, spacing, cones)
        d_dose = DeviceArray(shape, np.float32)
        if physics:
            _engine.convolve_rad(plan, d_terma, d_rho, d_dose, d_box,
                                 stream=s)
        elif use_fft:
            _engine.convolve_fft(plan, d_terma, d_dose, stream=s)
        else:
            _engine.convolve(plan, d_terma, d_dose, d_box, stream=s)
        dose = d_dose.download(stream=s)
        self.dose_grid_terma = d_terma.download(stream=s)
        self.dose_grid_densities = dens
        self.dose_grid.dose = dose
        s.sync()
        return dose
