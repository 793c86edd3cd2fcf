"""OpenLB checkpoint files: the text wrapper around a lattice's serialized bytes.

OpenLB writes ``Serializable::save(fileName)`` as one text file named
``<fileName>_rank<r>_size<n>.dat`` (src/core/serializer.hh:81-97, 131-137;
io/fileName.hh:51-64): the base64 encoding of the byte count (a size_t, or an
unsigned int with ``enforceUint``), followed by the base64 encoding of the
bytes, each with its own '=' padding (io/serializerIO.hh:40-64, io/base64.hh
with the standard alphabet).  The bytes themselves come from
``lbmh_checkpoint_save`` (one block) and are the same ones OpenLB's
``Serializable::save(std::uint8_t*)`` produces.
"""
from __future__ import annotations

import base64
import struct


def parallel_file_name(name: str, rank: int = 0, size: int = 1) -> str:
    """createParallelFileName(name) + ".dat" of a run with `size` ranks."""
    return f"{name}_rank{rank}_size{size}.dat"


def encode_dat(raw: bytes, enforce_uint: bool = False) -> str:
    if enforce_uint:
        if len(raw) > 0xFFFFFFFF:
            raise ValueError("checkpoint too large for a 32-bit size")
        head = struct.pack("<I", len(raw))
    else:
        head = struct.pack("<Q", len(raw))
    return base64.b64encode(head).decode("ascii") + base64.b64encode(raw).decode("ascii")


def decode_dat(text: str, enforce_uint: bool = False) -> bytes:
    text = "".join(text.split())   # OpenLB's decoder skips white space
    head = 8 if enforce_uint else 12
    if len(text) < head:
        raise ValueError("not an OpenLB checkpoint file")
    size = struct.unpack("<I" if enforce_uint else "<Q", base64.b64decode(text[:head]))[0]
    body = base64.b64decode(text[head:])
    if len(body) < size:
        raise ValueError(f"checkpoint file holds {len(body)} bytes, its header says {size}")
    return body[:size]


def write_file(path: str, raw: bytes, enforce_uint: bool = False) -> None:
    with open(path, "w") as fh:
        fh.write(encode_dat(raw, enforce_uint))


def read_file(path: str, enforce_uint: bool = False) -> bytes:
    with open(path) as fh:
        return decode_dat(fh.read(), enforce_uint)
