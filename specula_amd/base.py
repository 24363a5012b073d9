"""Minimal re-creation of SPECULA's processing-object contract, on torch device tensors.

Mirrors (names, arguments, lifecycle and error behaviour):
  BaseTimeObj        specula/base_time_obj.py:6-87   (device/precision selection, integer time)
  BaseDataObj        specula/base_data_obj.py:18-106 (generation_time, copyTo/transferDataTo)
  BaseValue          specula/base_value.py:5-58
  InputValue/List    specula/connections.py:2-80
  BaseProcessingObj  specula/base_processing_obj.py:8-184 (setup -> check_ready ->
                     prepare_trigger -> trigger -> post_trigger -> finalize; optional graph)

Arrays are torch tensors on ``cuda:<target_device_idx>`` (or the CPU for index -1).
"""
import torch

import specula_amd


def torch_device(target_device_idx):
    return torch.device('cpu') if target_device_idx < 0 else torch.device(f'cuda:{target_device_idx}')


class BaseTimeObj:
    """specula/base_time_obj.py:6-87."""

    def __init__(self, target_device_idx=None, precision=None):
        self._time_resolution = int(1e9)
        self.precision = specula_amd.global_precision if precision is None else precision
        self.target_device_idx = (specula_amd.default_target_device_idx
                                  if target_device_idx is None else target_device_idx)
        self.device = torch_device(self.target_device_idx)
        self.dtype = torch.float64 if self.precision == 0 else torch.float32
        self.complex_dtype = torch.complex128 if self.precision == 0 else torch.complex64

    def t_to_seconds(self, t):
        return float(t) / float(self._time_resolution)

    def seconds_to_t(self, seconds):
        """Exact decimal conversion, base_time_obj.py:74-87."""
        if self._time_resolution == 0:
            return 0
        ss = f'{float(seconds):.9f}'.rstrip('0').rstrip('.')
        if '.' not in ss:
            ss += '.0'
        intpart, fracpart = ss.split('.')
        return (int(intpart) * self._time_resolution +
                int(fracpart) * (self._time_resolution // (10 ** len(fracpart))))


class BaseDataObj(BaseTimeObj):
    """specula/base_data_obj.py:18-106: generation time + cross-device copies."""

    def __init__(self, target_device_idx=None, precision=None):
        super().__init__(target_device_idx, precision)
        self._generation_time = -1

    @property
    def generation_time(self):
        return self._generation_time

    @generation_time.setter
    def generation_time(self, value):
        self._generation_time = value

    def _tensor_attrs(self):
        return [k for k, v in vars(self).items() if isinstance(v, torch.Tensor)]

    def copyTo(self, target_device_idx):
        """Return a copy of this object on another device (base_data_obj.py:77-106)."""
        import copy
        if target_device_idx == self.target_device_idx:
            return self
        cloned = copy.copy(self)
        cloned.target_device_idx = target_device_idx
        cloned.device = torch_device(target_device_idx)
        for k in self._tensor_attrs():
            setattr(cloned, k, getattr(self, k).to(cloned.device, copy=True))
        return cloned

    def transferDataTo(self, destobj):
        """Copy array data into an existing object on another device (base_data_obj.py:58-75)."""
        for k in self._tensor_attrs():
            dst = getattr(destobj, k, None)
            src = getattr(self, k)
            if isinstance(dst, torch.Tensor) and dst.shape == src.shape:
                dst.copy_(src)
            else:
                setattr(destobj, k, src.to(destobj.device, copy=True))
        for k, v in vars(self).items():
            if not isinstance(v, torch.Tensor) and k not in ('device', 'target_device_idx'):
                setattr(destobj, k, v)
        return destobj


class BaseValue(BaseDataObj):
    """specula/base_value.py:5-58."""

    def __init__(self, description='', value=None, target_device_idx=None):
        super().__init__(target_device_idx=target_device_idx)
        self._description = description
        self._value = value

    def set_value(self, val):
        self._value = val

    @property
    def value(self):
        return self._value

    @value.setter
    def value(self, val):
        self._value = val

    @property
    def ptr_value(self):
        return self._value

    @ptr_value.setter
    def ptr_value(self, val):
        self._value = val


class InputValue:
    """specula/connections.py:2-33."""

    def __init__(self, type, optional=False):
        self.wrapped_type = type
        self.wrapped_value = None
        self.cloned_value = None
        self.optional = optional

    def get_time(self):
        if self.wrapped_value is not None:
            return self.wrapped_value.generation_time

    def get(self, target_device_idx):
        if self.wrapped_value is not None:
            if self.wrapped_value.target_device_idx == target_device_idx:
                return self.wrapped_value
            if self.cloned_value is None:
                self.cloned_value = self.wrapped_value.copyTo(target_device_idx)
            else:
                self.wrapped_value.transferDataTo(self.cloned_value)
            return self.cloned_value

    def set(self, value):
        if not isinstance(value, self.wrapped_type):
            raise ValueError(f'Value must be of type {self.wrapped_type} instead of {type(value)}')
        self.wrapped_value = value

    def type(self):
        return self.wrapped_type


class InputList:
    """specula/connections.py:36-80."""

    def __init__(self, type, optional=False):
        self.wrapped_type = type
        self.wrapped_list = None
        self.cloned_list = []
        self.optional = optional

    def get_time(self):
        if self.wrapped_list is not None:
            return [x.generation_time for x in self.wrapped_list]
        return []

    def get(self, target_device_idx):
        if self.wrapped_list is None:
            return None
        if self.cloned_list == []:
            for w in self.wrapped_list:
                self.cloned_list.append(w if w.target_device_idx == target_device_idx
                                        else w.copyTo(target_device_idx))
        else:
            for i, (w, c) in enumerate(zip(self.wrapped_list, self.cloned_list)):
                if w.target_device_idx == target_device_idx:
                    self.cloned_list[i] = w
                else:
                    w.transferDataTo(c)
        return self.cloned_list

    def set(self, new_list):
        for value in new_list:
            if not isinstance(value, self.wrapped_type):
                raise ValueError(f'List element must be of type {self.wrapped_type}')
        self.wrapped_list = new_list

    def type(self):
        return self.wrapped_type


class BaseProcessingObj(BaseTimeObj):
    """specula/base_processing_obj.py:8-184.

    ``trigger()`` replays a captured HIP graph of ``trigger_code()`` when the object has one
    (``build_stream(capture=True)``), else calls ``trigger_code()`` on the current stream."""

    _streams = {}
    # smallest captured graph that replaces its launches (see capture_stream)
    graph_min_nodes = 16

    def __init__(self, target_device_idx=None, precision=None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.current_time = 0
        self.current_time_seconds = 0
        self._verbose = 0
        self.stream = None
        self.ready = False
        self.cuda_graph = None
        self.graph_nodes = 0
        self.inputs = {}
        self.local_inputs = {}
        self.last_seen = {}
        self.outputs = {}

    def checkInputTimes(self):
        if len(self.inputs) == 0:
            return True
        for name, inp in self.inputs.items():
            if type(inp) is InputValue:
                t = inp.get_time()
                if name not in self.last_seen and t is not None and t >= 0:
                    return True
                if name in self.last_seen and t is not None and t > self.last_seen[name]:
                    return True
            elif type(inp) is InputList:
                if name not in self.last_seen:
                    if any(tt >= 0 for tt in inp.get_time()):
                        return True
                else:
                    for tt, last in zip(inp.get_time(), self.last_seen[name]):
                        if tt > last:
                            return True
        return False

    def prepare_trigger(self, t):
        self.current_time_seconds = self.t_to_seconds(self.current_time)
        for name, inp in self.inputs.items():
            if type(inp) is InputValue:
                self.local_inputs[name] = inp.get(self.target_device_idx)
                if self.local_inputs[name] is not None:
                    self.last_seen[name] = self.local_inputs[name].generation_time
            elif type(inp) is InputList:
                self.local_inputs[name] = []
                self.last_seen[name] = []
                lst = inp.get(self.target_device_idx)
                if lst is not None:
                    for tt in lst:
                        self.local_inputs[name].append(tt)
                        self.last_seen[name].append(tt.generation_time)

    def trigger_code(self):
        pass

    def post_trigger(self):
        # base_processing_obj.py:93-96 synchronises the object's stream after a graph launch so
        # that the next object sees the outputs.  Here graphs are replayed, and eager triggers
        # run, on the caller's current stream, so later work on it is ordered after these outputs
        # without any synchronisation.
        pass

    @classmethod
    def device_stream(cls, target_device_idx):
        if target_device_idx not in cls._streams:
            with torch.cuda.device(target_device_idx):
                cls._streams[target_device_idx] = torch.cuda.Stream()   # default priority
        return cls._streams[target_device_idx]

    def build_stream(self, allow_parallel=True, capture=False):
        """Create (or share) the object's stream; optionally capture trigger_code() in a
        HIP graph after one eager warm-up run, as base_processing_obj.py:114-131."""
        if self.target_device_idx < 0:
            return
        with torch.cuda.device(self.target_device_idx):
            self.stream = torch.cuda.Stream() if allow_parallel else self.device_stream(self.target_device_idx)
            if capture:
                self.capture_stream()

    def capture_stream(self):
        """One eager warm-up of trigger_code() on the object's stream, then its library calls
        captured into a HIP graph by the native library (shwfs_graph_*).  The graph is kept only
        when it holds at least ``graph_min_nodes`` launches: on ROCm a graph launch costs the host
        ~15 us and slows the next direct launch on the stream by ~12 us, against ~2 us per direct
        kernel launch, so a graph of a few launches replays slower than the launches themselves.
        A declined capture leaves ``cuda_graph`` None and ``trigger()`` runs trigger_code()."""
        from specula_amd import _native
        with torch.cuda.device(self.target_device_idx):
            self.stream.wait_stream(torch.cuda.current_stream())     # the warm-up reads the inputs
            with torch.cuda.stream(self.stream):
                self.trigger_code()
            self.stream.synchronize()
            graph = _native.HipGraph().capture(self.trigger_code, self.stream)
            self.graph_nodes = graph.node_count()
            if self.graph_nodes < self.graph_min_nodes:
                graph.destroy()
                self.cuda_graph = None
            else:
                self.cuda_graph = graph

    def check_ready(self, t):
        self.current_time = t
        if self.checkInputTimes():
            self.prepare_trigger(t)
            self.ready = True
        elif self.verbose:
            print('No inputs have been refreshed, skipping trigger')
        return self.ready

    def trigger(self):
        """base_processing_obj.py:145-154: a captured graph is launched, else trigger_code() runs
        (native calls take stream=None: the current stream).  Both go on the caller's current
        stream, which produced the inputs; the object's own stream only serves the capture."""
        if self.ready:
            if self.target_device_idx >= 0:
                with torch.cuda.device(self.target_device_idx):
                    if self.cuda_graph is not None:
                        self.cuda_graph.replay()
                    else:
                        self.trigger_code()
            else:
                self.trigger_code()
            self.ready = False

    @property
    def verbose(self):
        return self._verbose

    @verbose.setter
    def verbose(self, value):
        self._verbose = value

    def setup(self):
        """Check that every mandatory input is set (base_processing_obj.py:164-177)."""
        for name, inp in self.inputs.items():
            if inp.get(self.target_device_idx) is None and not inp.optional:
                raise ValueError(f'Input {name} for object {self} has not been set')

    def finalize(self):
        pass
