"""Live field animation during simulation (drop-in for ``pyfds.gfx``, ``pyfds/gfx.py:1-440``).

The reference runs the simulation in a forked process and sends every frame (and finally the
whole field) back through a multiprocessing queue (``pyfds/gfx.py:71-98``).  A process that
holds GPU contexts cannot fork, and on the device the simulation no longer needs a process of
its own: here a frame is one device run of ``steps_per_frame`` steps (graph-replayed chunks)
followed by one device-to-host copy of the observed component, produced in the plotting
process by :meth:`Animator.frames`.  The field's components, outputs and step counter are
therefore the field's own after the animation, with nothing to copy back.

Plot layout, labels, region drawing, frame files and the ffmpeg call keep the reference's
behaviour.
"""
import os

import matplotlib.patches as pa
import matplotlib.pyplot as pp
import numpy as np

from . import fields as fld
from . import regions as reg

__all__ = [
    'Animator1D', 'Animator2D',
]


class Animator:
    """Base class of the live animation (``pyfds/gfx.py:14-118``)."""

    def __init__(self, field, observed_component=None, steps_per_frame=10, scale=1,
                 frame_delay=1e-2, save_video=False, video_file_name='pyfds_vid.mp4'):
        self.field = field
        self.field_components = {name: getattr(self.field, name) for name in dir(self.field)
                                 if isinstance(getattr(self.field, name), fld.FieldComponent)}
        if observed_component:
            if observed_component not in self.field_components:
                raise KeyError('Field component {} not found in given field.'
                               .format(observed_component))
            self.observed_component = observed_component
        else:
            self.observed_component = list(self.field_components.keys())[0]

        self.steps_per_frame = int(steps_per_frame)
        if isinstance(scale, (list, tuple, np.ndarray)) and len(scale) == 2:
            self.scale = scale
        elif isinstance(scale, (float, int)):
            self.scale = (-scale, scale)
        else:
            raise ValueError('Scale must either be scalar or a 2 element vector.')
        self.frame_delay = frame_delay
        self.save_video = save_video
        self.video_file_name = video_file_name
        self.image_files = []
        self.ffmpeg_path = 'ffmpeg'
        self.video_fps = 30
        self.video_dpi = 100

        self.show_boundaries = True
        self.show_materials = True
        self.show_output = True

        self._x_axis_prefix, self._x_axis_factor = get_prefix(max(self.field.x.vector))
        self._t_prefix, self._t_factor = get_prefix(max(self.field.t.vector))

        self.axes = None
        self.plot_title = ''
        self.x_label = '$x$'
        self.time_precision = 2

    # -- frames --------------------------------------------------------------------------
    def frames(self):
        """Yield ``(time, values)`` of the observed component after every ``steps_per_frame``
        steps, ``t.samples // steps_per_frame`` frames in all (the reference's
        ``_sim_function`` loop, ``pyfds/gfx.py:71-86``)."""
        for _ in range(int(self.field.t.samples / self.steps_per_frame)):
            self.field.simulate(self.steps_per_frame)
            values = np.array(getattr(self.field, self.observed_component).values)
            yield self.field.t.vector[self.field.step - 1], values

    def _show_frames(self, set_data):
        """Draw every frame as it arrives (``pyfds/gfx.py:220-246``)."""
        for time, data in self.frames():
            self.axes.title.set_text('{title} $t$ = {time:.{prec}f} {prefix}s'
                                     .format(title=self.plot_title, time=time / self._t_factor,
                                             prec=self.time_precision, prefix=self._t_prefix))
            set_data(data)
            pp.pause(self.frame_delay)
            if self.save_video:
                self._save_frame()
        if self.save_video:
            self._create_video()

    # -- video -----------------------------------------------------------------------------
    def _save_frame(self):
        """Save the current frame as a png file for the video."""
        file_name = 'frame{0:06d}.png'.format(len(self.image_files))
        pp.savefig(file_name, dpi=self.video_dpi)
        self.image_files.append(file_name)

    def _create_video(self):
        """Join the saved frames into ``video_file_name`` with ffmpeg, asking before
        overwriting, and delete the frame files (``pyfds/gfx.py:96-118``)."""
        answer = None
        if os.path.isfile(self.video_file_name):
            answer = input('File {} already exists. Overwrite? [y/N]'
                           .format(self.video_file_name))
            if answer.capitalize() == 'Y':
                os.remove(self.video_file_name)
        if not answer or answer.capitalize() == 'Y':
            # the codec needs even image sizes; 6.4 x 4.78 inch is matplotlib's default figure
            width = int((self.video_dpi * 6.4) // 2 * 2)
            height = int((self.video_dpi * 4.78) // 2 * 2)
            os.system('{} -loglevel error -framerate {} -i frame%06d.png -pix_fmt yuv420p -vf '
                      'scale={}:{} {}'.format(self.ffmpeg_path, self.video_fps, width, height,
                                              self.video_file_name))
        for file in self.image_files:
            os.remove(file)

    # -- setup plot --------------------------------------------------------------------------
    def _plot_regions(self):
        if self.show_materials:
            for mat_region in self.field.material_regions:
                self.plot_region(mat_region.region)
        for flag, kind in ((self.show_boundaries, 'boundaries'), (self.show_output, 'outputs')):
            if flag:
                for component in self.field_components.values():
                    for item in getattr(component, kind):
                        self.plot_region(item.region)


class Animator1D(Animator):
    """Animation of a 1D field (``pyfds/gfx.py:121-249``)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.y_label = self.observed_component

    def plot_region(self, region):
        """Mark a point or line region with vertical lines."""
        if isinstance(region, reg.PointRegion):
            self.axes.plot(np.ones(2) * region.point_coordinates / self._x_axis_factor,
                           self.scale, color='black')
        elif isinstance(region, reg.LineRegion):
            for edge in region.line_coordinates[:2]:
                self.axes.plot(np.ones(2) * edge / self._x_axis_factor, self.scale,
                               color='black')
        else:
            raise TypeError('Unknown type in region list: {}'.format(type(region)))

    def show_setup(self, halt=True):
        """Plot the setup: axes, materials, boundaries and outputs."""
        pp.figure()
        self.axes = pp.gca()
        self.axes.set_xlim(0, max(self.field.x.vector) / self._x_axis_factor)
        self.axes.set_ylim(self.scale)
        self.axes.set_xlabel('{0} / {1}m'.format(self.x_label, self._x_axis_prefix))
        self.axes.set_ylabel(self.y_label)
        pp.grid(True)
        self._plot_regions()
        if halt:
            pp.show()

    def start_simulation(self):
        """Run the simulation and animate the observed component."""
        self.show_setup(halt=False)
        main_plot, = self.axes.plot([])
        x = self.field.x.vector / self._x_axis_factor
        if self.field.t.samples == 1:
            self.field.simulate()
            main_plot.set_data(x, getattr(self.field, self.observed_component).values)
        else:
            self._show_frames(lambda data: main_plot.set_data(x, data))
        pp.show()


class Animator2D(Animator):
    """Animation of a 2D field (``pyfds/gfx.py:252-418``)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._y_axis_prefix = self._x_axis_prefix
        self._y_axis_factor = self._x_axis_factor
        self.y_label = '$y$'
        self.c_label = self.observed_component

    def plot_region(self, region):
        """Outline a point, line, rectangle, triangle or ellipse region."""
        fx, fy = self._x_axis_factor, self._y_axis_factor
        if isinstance(region, reg.PointRegion):
            self.axes.scatter(region.point_coordinates[0] / fx, region.point_coordinates[1] / fy,
                              color='black')
        elif isinstance(region, reg.LineRegion):
            c = region.line_coordinates
            self.axes.plot([c[0] / fx, c[2] / fx], [c[1] / fy, c[3] / fy], color='black')
        elif isinstance(region, reg.RectRegion):
            c = region.rect_coordinates
            self.axes.add_patch(pa.Rectangle((c[0] / fx, c[1] / fy), c[2] / fx, c[3] / fy,
                                             fill=False))
        elif isinstance(region, reg.TriRegion):
            c = region.tri_coordinates
            self.axes.add_patch(pa.Polygon(np.array([[c[0] / fx, c[1] / fy],
                                                     [c[2] / fx, c[3] / fy],
                                                     [c[4] / fx, c[5] / fy]]), fill=False))
        elif isinstance(region, reg.EllipseRegion):
            self.axes.add_patch(pa.Ellipse((region.centre[0] / fx, region.centre[1] / fy),
                                           2 * region.radii[0] / fx, 2 * region.radii[1] / fy,
                                           fill=False))
        else:
            raise TypeError('Unknown type in region list: {}'.format(type(region)))

    def field_as_matrix(self, component=None):
        """The observed component (or ``component``) as a (y, x) matrix."""
        if component is None:
            component = getattr(self.field, self.observed_component).values
        return np.reshape(component, (self.field.y.samples, self.field.x.samples))

    def show_setup(self, halt=True):
        """Plot the setup: axes, materials, boundaries and outputs."""
        pp.figure()
        self.axes = pp.gca()
        pp.axis('equal')
        self.axes.set_xlim(0, max(self.field.x.vector) / self._x_axis_factor)
        self.axes.set_ylim(0, max(self.field.y.vector) / self._y_axis_factor)
        self.axes.set_xlabel('{0} / {1}m'.format(self.x_label, self._x_axis_prefix))
        self.axes.set_ylabel('{0} / {1}m'.format(self.y_label, self._y_axis_prefix))
        self._plot_regions()
        if halt:
            pp.show()

    def start_simulation(self):
        """Run the simulation and animate the observed component."""
        self.show_setup(halt=False)
        main_plot = self.axes.imshow(self.field_as_matrix(),
                                     extent=(0, max(self.field.x.vector) / self._x_axis_factor,
                                             max(self.field.y.vector) / self._y_axis_factor, 0),
                                     cmap='viridis')
        main_plot.set_clim(self.scale)
        color_bar = pp.colorbar(main_plot)
        color_bar.set_label(self.observed_component, rotation=270)
        if self.field.t.samples == 1:
            self.field.simulate()
            main_plot.set_data(self.field_as_matrix())
        else:
            self._show_frames(lambda data: main_plot.set_data(self.field_as_matrix(data)))
        pp.show()


def get_prefix(value):
    """Metric prefix and factor for ``value`` (``pyfds/gfx.py:421-435``)."""
    for factor, prefix in sorted(prefixes.items()):
        if value / factor < 1e3:
            break
    return prefix, factor


prefixes = {
    1e-24: 'y', 1e-21: 'z', 1e-18: 'a', 1e-15: 'f', 1e-12: 'p', 1e-9: 'n', 1e-6: 'µ',
    1e-3: 'm', 1e0: '', 1e3: 'k', 1e6: 'M', 1e9: 'G', 1e12: 'T', 1e15: 'P', 1e18: 'E',
    1e21: 'Z', 1e24: 'Y',
}
