"""CNN towers (jax_cfd/ml/towers.py:144-262) as descriptions the GPU runs.

``forward_tower_factory`` returns a ConvTower: the shapes of its periodic
3x3 convolutions (relu between them, identity input/output scaling) and a
seeded initialiser with haiku Conv2D's defaults.  The tower itself runs in
libjcfd_amd (li_conv3x3, matrix-core implicit GEMM); forms the kernels do
not implement (other kernel sizes, dilations, strides, scale functions,
nonlinearities) raise NotImplementedError here.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Tuple, Union

import numpy as np

HIDDEN_CHANNELS_BUILT = (16, 32, 64)


@dataclasses.dataclass(frozen=True)
class ConvTower:
  """conv(2 -> hidden), relu, ..., conv(hidden -> num_output_channels)."""
  num_output_channels: int
  ndim: int
  num_hidden_channels: int = 16
  num_hidden_layers: int = 2
  num_inputs: int = 2
  name: str = 'forward_cnn_tower'

  @property
  def shapes(self) -> List[Tuple[int, int, int, int]]:
    """(kh, kw, cin, cout) of each convolution, haiku's Conv2D layout."""
    out, cin = [], self.num_inputs
    for _ in range(self.num_hidden_layers):
      out.append((3, 3, cin, self.num_hidden_channels))
      cin = self.num_hidden_channels
    out.append((3, 3, cin, self.num_output_channels))
    return out

  def init_params(self, seed: int = 0, dtype=np.float32):
    """[(w, b)] with haiku Conv2D's default initialiser: w from a truncated
    normal (2 standard deviations) of stddev sqrt(1 / fan_in) / 0.8796...,
    fan_in = kh kw cin; b = 0."""
    rng = np.random.default_rng(seed)
    params = []
    for kh, kw, cin, cout in self.shapes:
      std = np.sqrt(1.0 / (kh * kw * cin)) / 0.87962566103423978
      w = rng.standard_normal((kh, kw, cin, cout))
      out = np.abs(w) > 2
      while out.any():
        w[out] = rng.standard_normal(int(out.sum()))
        out = np.abs(w) > 2
      params.append(((w * std).astype(dtype), np.zeros(cout, dtype)))
    return params

  def check_params(self, params):
    params = list(params)
    if len(params) != len(self.shapes):
      raise ValueError(f'{self.name}: expected {len(self.shapes)} layers, got {len(params)}')
    out = []
    for k, ((w, b), shape) in enumerate(zip(params, self.shapes)):
      w = np.ascontiguousarray(np.asarray(w, np.float32))
      b = np.ascontiguousarray(np.asarray(b, np.float32))
      if w.shape != shape or b.shape != (shape[3],):
        raise ValueError(f'{self.name}: layer {k} has w {w.shape}, b {b.shape}; '
                         f'expected {shape}, ({shape[3]},)')
      out.append((w, b))
    return out


def _unsupported(what):
  raise NotImplementedError(f'forward_tower_factory: {what} is not built '
                            '(3x3 periodic convolutions, rate 1, stride 1, relu, identity '
                            'scaling, 16/32/64 hidden channels)')


def forward_tower_factory(num_output_channels: int, ndim: int, num_hidden_channels: int = 16,
                          kernel_size: int = 3, num_hidden_layers: int = 2,
                          rates: Union[int, Tuple[int, ...]] = 1,
                          strides: Union[int, Tuple[int, ...]] = 1,
                          output_kernel_size: int = 3, output_dilation_rate: int = 1,
                          output_stride: int = 1, conv_module=None, nonlinearity=None,
                          inputs_scale_fn=None, output_scale_fn=None,
                          name: Optional[str] = 'forward_cnn_tower') -> ConvTower:
  """towers.py:144-198 with the built options."""
  if ndim != 2:
    _unsupported(f'ndim={ndim}')
  if kernel_size != 3 or output_kernel_size != 3:
    _unsupported('a kernel size other than 3')
  if any(np.atleast_1d(rates) != 1) or output_dilation_rate != 1:
    _unsupported('dilation')
  if any(np.atleast_1d(strides) != 1) or output_stride != 1:
    _unsupported('stride')
  for fn, what in ((conv_module, 'conv_module'), (nonlinearity, 'nonlinearity'),
                   (inputs_scale_fn, 'inputs_scale_fn'), (output_scale_fn, 'output_scale_fn')):
    if fn is not None and getattr(fn, '__name__', '') not in ('periodic_convolution', 'relu'):
      _unsupported(what)
  if num_hidden_layers < 0 or num_hidden_layers + 1 > 16:
    _unsupported(f'{num_hidden_layers} hidden layers')
  if num_hidden_layers > 0 and num_hidden_channels not in HIDDEN_CHANNELS_BUILT:
    _unsupported(f'{num_hidden_channels} hidden channels')
  return ConvTower(int(num_output_channels), ndim, int(num_hidden_channels),
                   int(num_hidden_layers), name=name or 'forward_cnn_tower')
