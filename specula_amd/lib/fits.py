"""Minimal FITS image reader/writer (primary + IMAGE extensions), no astropy.

Enough for the files on the SH path: SubapData (subap_data.py:53-79), Slopes
(slopes.py:125-156), ConvolutionKernel v1.1 (convolution_kernel.py:276-305) and the reference
test fixtures.  Data are big-endian per the FITS standard; BZERO/BSCALE are applied.
"""
import numpy as np

_BLOCK = 2880
_BITPIX = {8: '>u1', 16: '>i2', 32: '>i4', 64: '>i8', -32: '>f4', -64: '>f8'}


def _parse_value(v):
    v = v.strip()
    if v.startswith("'"):
        end = v.find("'", 1)
        while end != -1 and end + 1 < len(v) and v[end + 1] == "'":
            end = v.find("'", end + 2)
        return v[1:end].replace("''", "'").rstrip()
    if '/' in v:
        v = v.split('/')[0].strip()
    if v == 'T':
        return True
    if v == 'F':
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace('D', 'E'))
    except ValueError:
        return v


def _read_header(buf, pos):
    hdr = {}
    while True:
        block = buf[pos:pos + _BLOCK]
        if len(block) < _BLOCK:
            raise ValueError('truncated FITS header')
        pos += _BLOCK
        done = False
        for i in range(0, _BLOCK, 80):
            card = block[i:i + 80].decode('ascii', errors='replace')
            key = card[:8].strip()
            if key == 'END':
                done = True
                break
            if card[8:10] == '= ':
                hdr[key] = _parse_value(card[10:])
        if done:
            return hdr, pos


def read(filename):
    """Return a list of (data or None, header dict) for every HDU."""
    with open(filename, 'rb') as fh:
        buf = fh.read()
    pos = 0
    hdus = []
    while pos < len(buf):
        if not buf[pos:pos + 80].strip():
            break
        hdr, pos = _read_header(buf, pos)
        naxis = hdr.get('NAXIS', 0)
        shape = tuple(hdr[f'NAXIS{k}'] for k in range(naxis, 0, -1))
        nbytes = 0
        data = None
        if naxis > 0 and all(s > 0 for s in shape):
            dt = np.dtype(_BITPIX[hdr['BITPIX']])
            count = int(np.prod(shape))
            nbytes = count * dt.itemsize
            data = np.frombuffer(buf, dtype=dt, count=count, offset=pos).reshape(shape)
            data = data.astype(dt.newbyteorder('='))
            bscale, bzero = hdr.get('BSCALE', 1), hdr.get('BZERO', 0)
            if bscale != 1 or bzero != 0:
                if hdr['BITPIX'] == 64 and bzero == 2 ** 63:
                    data = (data.astype(np.int64).view(np.uint64) ^ np.uint64(2 ** 63))
                else:
                    data = data * bscale + bzero
        pos += (nbytes + _BLOCK - 1) // _BLOCK * _BLOCK
        hdus.append((data, hdr))
    return hdus


def _card(key, value):
    if isinstance(value, bool):
        v = 'T' if value else 'F'
        s = f'{key:<8}= {v:>20}'
    elif isinstance(value, (int, np.integer)):
        s = f'{key:<8}= {int(value):>20}'
    elif isinstance(value, (float, np.floating)):
        s = f'{key:<8}= {repr(float(value)).upper():>20}'
    else:
        v = "'" + str(value).replace("'", "''") + "'"
        s = f'{key:<8}= {v:<20}'
    return s[:80].ljust(80)


def write(filename, hdus):
    """hdus: list of (array or None, header dict); the first becomes the primary HDU."""
    out = bytearray()
    for k, (data, hdr) in enumerate(hdus):
        cards = []
        if k == 0:
            cards.append(_card('SIMPLE', True))
        else:
            cards.append(_card('XTENSION', 'IMAGE'))
        arr = None if data is None else np.asarray(data)
        if arr is not None and arr.dtype.kind == 'b':
            arr = arr.astype(np.uint8)
        if arr is not None:
            code = {np.dtype(np.uint8): 8, np.dtype(np.int16): 16, np.dtype(np.int32): 32,
                    np.dtype(np.int64): 64, np.dtype(np.float32): -32, np.dtype(np.float64): -64}
            bitpix = code.get(arr.dtype)
            if bitpix is None:
                arr = arr.astype(np.float64)
                bitpix = -64
        else:
            bitpix = 8
        cards.append(_card('BITPIX', bitpix))
        shape = () if arr is None else arr.shape
        cards.append(_card('NAXIS', len(shape)))
        for i, s in enumerate(reversed(shape)):
            cards.append(_card(f'NAXIS{i + 1}', s))
        if k == 0:
            cards.append(_card('EXTEND', True))
        else:
            cards.append(_card('PCOUNT', 0))
            cards.append(_card('GCOUNT', 1))
        for key, value in hdr.items():
            cards.append(_card(key, value))
        cards.append('END'.ljust(80))
        head = ''.join(cards).encode('ascii')
        out += head + b' ' * ((-len(head)) % _BLOCK)
        if arr is not None:
            raw = np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder('>')).tobytes()
            out += raw + b'\0' * ((-len(raw)) % _BLOCK)
    with open(filename, 'wb') as fh:
        fh.write(out)
