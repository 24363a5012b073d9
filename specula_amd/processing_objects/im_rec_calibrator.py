"""Drop-in ImRecCalibrator, specula/processing_objects/im_rec_calibrator.py:10-98.

Accumulates the push-pull interaction matrix on the device, one row update per step
(``IM[mode] += slopes / command`` for the step's non-zero command), and at ``finalize`` writes the
Intmat / Recmat FITS files under ``data_dir``.
"""
import os

import torch

from specula_amd.base import BaseProcessingObj, BaseValue, InputValue
from specula_amd.data_objects.intmat import Intmat
from specula_amd.data_objects.slopes import Slopes


class ImRecCalibrator(BaseProcessingObj):
    def __init__(self, nmodes: int, data_dir: str, rec_tag: str, im_tag: str = '', first_mode: int = 0,
                 pupdata_tag: str = None, tag_template: str = None,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self._nmodes = nmodes
        self._first_mode = first_mode
        self._data_dir = data_dir
        if tag_template is None and (rec_tag is None or rec_tag == 'auto'):
            raise ValueError('At least one of tag_template and rec_tag must be set')
        self._rec_filename = tag_template if (rec_tag is None or rec_tag == 'auto') else rec_tag
        self._im_filename = im_tag
        self._im = None
        self.inputs['in_slopes'] = InputValue(type=Slopes)
        self.inputs['in_commands'] = InputValue(type=BaseValue)
        self.output_im = [Slopes(length=2, target_device_idx=self.target_device_idx) for _ in range(nmodes)]
        self.outputs['out_im'] = self.output_im
        self.pupdata_tag = pupdata_tag

    @property
    def im(self):
        return self._im

    def trigger_code(self):
        """im_rec_calibrator.py:42-71."""
        sl = self.local_inputs['in_slopes']
        if sl.generation_time != self.current_time:
            return
        slopes = sl.slopes
        commands = torch.as_tensor(self.local_inputs['in_commands'].value, device=self.device)
        if self._im is None:
            self._im = torch.zeros((self._nmodes, slopes.numel()), dtype=self.dtype, device=self.device)
            for o in self.output_im:
                o.resize(slopes.numel())
        nz = torch.nonzero(commands).flatten().tolist()
        if len(nz) > 1:
            raise ValueError('push-pull commands must have one non-zero element per step')
        if nz:
            mode = nz[0] - self._first_mode
            if mode < self._nmodes:
                self._im[mode] += slopes.to(self.dtype) / commands[nz[0]].to(self.dtype)
        for i, o in enumerate(self.output_im):
            o.slopes[:] = self._im[i]
            o.single_mask = sl.single_mask
            o.display_map = sl.display_map
            o.generation_time = self.current_time

    def finalize(self):
        im = Intmat(self._im, pupdata_tag=self.pupdata_tag, target_device_idx=self.target_device_idx,
                    precision=self.precision)
        if self._im_filename:
            im.save(os.path.join(self._data_dir, self._im_filename))
        if self._rec_filename:
            rec = im.generate_rec(self._nmodes)
            rec.save(os.path.join(self._data_dir, self._rec_filename))

    def setup(self):
        super().setup()
        for name in (self._im_filename, self._rec_filename):
            if name:
                path = os.path.join(self._data_dir, name)
                path = path if path.endswith('.fits') else path + '.fits'
                if os.path.exists(path):
                    raise FileExistsError(f'{path} already exists, please remove it')
