"""Regions, boundaries, outputs and material regions (drop-in for ``pyfds.regions``).

Same classes, constructor arguments and attributes as the reference
(``pyfds/regions.py:8-180``).  These are plain host-side data objects: during
``Field.simulate``/``sim_step`` their semantics are executed on the device (boundaries are
compiled into per-cell programs, outputs into gather lists, see ``pyfds_amd.device``).
``Boundary.apply`` stays available as the reference's array-level helper.
"""
import numpy as np

__all__ = [
    'Boundary', 'MaterialRegion', 'Output',
]


class Region:
    """Point indices plus metadata (``pyfds/regions.py:8-21``)."""

    def __init__(self, indices, name=''):
        self.indices = indices
        self.name = name


class PointRegion(Region):
    """Region of individual points (``pyfds/regions.py:24-37``)."""

    def __init__(self, indices, coordinates, name=''):
        super().__init__(indices, name)
        self.point_coordinates = coordinates


class LineRegion(Region):
    """Region along a line of points (``pyfds/regions.py:40-53``)."""

    def __init__(self, indices, coordinates, name=''):
        super().__init__(indices, name)
        self.line_coordinates = coordinates


class RectRegion(Region):
    """Rectangular block of points (``pyfds/regions.py:56-69``)."""

    def __init__(self, indices, coordinates, name=''):
        super().__init__(indices, name)
        self.rect_coordinates = coordinates


class TriRegion(Region):
    """Triangular set of points (``pyfds/regions.py:72-85``)."""

    def __init__(self, indices, coordinates, name=''):
        super().__init__(indices, name)
        self.tri_coordinates = coordinates


class EllipseRegion(Region):
    """Elliptic set of points (``pyfds/regions.py:88-103``)."""

    def __init__(self, indices, centre, radii, name=''):
        super().__init__(indices, name)
        self.centre = centre
        self.radii = radii


class Boundary:
    """Values written into a FieldComponent every step (``pyfds/regions.py:106-145``).

    ``value`` is a scalar, a signal (ndarray indexed by the step) or a list with one signal
    per region point; ``additive`` selects ``old + value`` over ``value``.  The value is
    kept by reference, as in the reference: later in-place changes to a signal are seen by
    the next ``simulate`` call.
    """

    def __init__(self, region, value=0, additive=False):
        self.region = region
        self.value = value
        self.additive = additive

    def apply(self, old_values, step):
        """New values for the boundary points given their old values (host helper; the
        device applies the same rule inside the step kernels)."""
        if np.ndim(self.value) == 0:
            return self.additive * old_values + self.value
        if isinstance(self.value, np.ndarray):
            return self.additive * old_values + self.value[step]
        return [self.additive * old_values[k] + sig[step] for k, sig in enumerate(self.value)]


class Output:
    """Points recorded every step (``pyfds/regions.py:148-165``); ``signals`` holds one list
    per point.

    Device runs hand over their samples as one (steps x points) block; the blocks are turned
    into the per-point lists when ``signals`` is next read (building the lists costs about as
    much as a short run).  A list taken from ``signals`` before a run therefore does not grow
    during the run: read ``signals`` again afterwards."""

    def __init__(self, region):
        self.region = region
        self._signals = []
        self._blocks = []

    @property
    def signals(self):
        if self._blocks:
            self._merge()
        return self._signals

    @signals.setter
    def signals(self, value):
        self._blocks = []
        self._signals = value

    def _record(self, block):
        """Queue the samples of a device run (steps x points, float64, owned by this output)."""
        self._blocks.append(block)

    def _merge(self):
        blocks, self._blocks = self._blocks, []
        cols = (blocks[0] if len(blocks) == 1 else np.concatenate(blocks, axis=0)).T
        if not self._signals:
            self._signals = [list(col) for col in cols]
        else:
            for sig, col in zip(self._signals, cols):
                sig.extend(col)

    @property
    def mean_signal(self):
        """Mean over the points of the region, per step."""
        return np.mean(np.asarray(self.signals), axis=0)


class MaterialRegion:
    """A material assigned to a region (``pyfds/regions.py:168-180``)."""

    def __init__(self, region, material):
        self.region = region
        self.materials = [material]
