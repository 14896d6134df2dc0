"""i-PI socket driver (the reference's socket-driven mode, ipisocket.F90:175-359 and
fortnet.F90:126-169,503-609) against a scripted i-PI server on a UNIX socket.

CPU: the message sequence (STATUS/READY, POSDATA, STATUS/HAVEDATA, GETFORCE/FORCEREADY, EXIT),
the cell and position layouts, the virial the reference sends, and the errors for a wrong atom
count or an unexpected header, with a stand-in evaluation.  GPU: a water box driven for several
steps gives the energies and forces of a direct evaluation of each geometry.
"""
import os
import socket
import struct
import threading

import numpy as np
import pytest

import fortnet_amd as fa
from fortnet_amd import ipi
from fortnet_amd import workloads as W


class Server(threading.Thread):
    """Scripted i-PI server: sends each geometry, collects what the driver returns.  ``path`` is a
    UNIX socket path, or None for an internet socket on 127.0.0.1 (port in ``self.address``)."""

    def __init__(self, path, frames, natom_sent=None, bad_header=None):
        super().__init__(daemon=True)
        self.path = path
        self.frames = frames  # list of (coords [n, 3], h [3, 3] with the vectors as columns)
        self.natom_sent = natom_sent
        self.bad_header = bad_header
        self.results = []
        self.replies = []
        self.error = None
        if path is None:
            self.lsock = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
            self.lsock.bind(("127.0.0.1", 0))
            self.address = ("127.0.0.1", self.lsock.getsockname()[1])
        else:
            self.lsock = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
            self.lsock.bind(path)
            self.address = path
        self.lsock.listen(1)

    def _recv(self, c, n):
        buf = b""
        while len(buf) < n:
            chunk = c.recv(n - len(buf))
            if not chunk:
                raise EOFError
            buf += chunk
        return buf

    def _hdr(self, c, text):
        c.sendall(text.encode().ljust(12))

    def run(self):
        try:
            c, _ = self.lsock.accept()
            with c:
                if self.bad_header:
                    self._hdr(c, self.bad_header)
                    return
                for coords, h in self.frames:
                    for _ in range(2):
                        self._hdr(c, "STATUS")
                        self.replies.append(self._recv(c, 12).decode().strip())
                    self._hdr(c, "POSDATA")
                    c.sendall(np.ascontiguousarray(h, np.float64).tobytes())
                    c.sendall(np.linalg.inv(h).astype(np.float64).tobytes())
                    n = len(coords) if self.natom_sent is None else self.natom_sent
                    c.sendall(struct.pack("i", n))
                    c.sendall(np.ascontiguousarray(coords, np.float64).tobytes())
                    self._hdr(c, "STATUS")
                    self.replies.append(self._recv(c, 12).decode().strip())
                    self._hdr(c, "GETFORCE")
                    self.replies.append(self._recv(c, 12).decode().strip())
                    (e,) = struct.unpack("d", self._recv(c, 8))
                    (n,) = struct.unpack("i", self._recv(c, 4))
                    f = np.frombuffer(self._recv(c, 24 * n), np.float64).reshape(n, 3)
                    vir = np.frombuffer(self._recv(c, 72), np.float64).reshape(3, 3)
                    (nx,) = struct.unpack("i", self._recv(c, 4))
                    self.results.append((e, f.copy(), vir.copy(), nx))
                self._hdr(c, "STATUS")
                self.replies.append(self._recv(c, 12).decode().strip())
                self._hdr(c, "EXIT")
        except Exception as exc:  # reported by the test
            self.error = exc
        finally:
            self.lsock.close()


def _frames(n, steps, seed=0):
    rng = np.random.default_rng(seed)
    h = np.array([[10.0, 0.5, 0.0], [0.0, 9.0, 0.3], [0.0, 0.0, 11.0]])  # vectors as columns
    base = rng.uniform(0, 9, size=(n, 3))
    return [(base + 0.05 * rng.normal(size=base.shape), h) for _ in range(steps)]


def _harmonic(geo):
    return float((geo.coords ** 2).sum()), -2.0 * geo.coords


def test_protocol_round_trip(tmp_path):
    path = str(tmp_path / "ipi.sock")
    frames = _frames(5, 3)
    srv = Server(path, frames)
    srv.start()
    seen = []

    def ev(geo):
        seen.append(geo)
        return _harmonic(geo)

    drv = ipi.IpiDriver(path, [14] * 5, evaluate=ev, timeout=30)
    assert drv.run(10) == 3  # stops at EXIT before the step limit
    drv.close()
    srv.join(30)
    assert srv.error is None, srv.error
    assert srv.replies == ["READY", "READY", "HAVEDATA", "FORCEREADY"] * 3 + ["READY"]
    for (coords, h), geo, (e, f, vir, nx) in zip(frames, seen, srv.results):
        assert np.array_equal(geo.coords, coords)
        assert np.array_equal(geo.latvecs, h.T)  # lattice vectors as rows
        e0, f0 = _harmonic(geo)
        assert e == e0 and np.array_equal(f, f0)
        assert np.allclose(vir, abs(np.linalg.det(h)))  # the reference's virial: ones x volume
        assert nx == 0


def test_internet_socket(tmp_path):
    """The reference's internet-socket mode (host and port, ipisocket.F90:55-84) on 127.0.0.1."""
    frames = _frames(3, 2, seed=4)
    srv = Server(None, frames)
    srv.start()
    drv = ipi.IpiDriver(srv.address, [8, 1, 1], evaluate=_harmonic, timeout=30)
    assert drv.run(5) == 2
    drv.close()
    srv.join(30)
    assert srv.error is None, srv.error
    for (coords, h), (e, f, vir, nx) in zip(frames, srv.results):
        assert e == float((coords ** 2).sum()) and np.array_equal(f, -2.0 * coords)


def test_step_limit_and_wrong_atom_count(tmp_path):
    path = str(tmp_path / "a.sock")
    srv = Server(path, _frames(4, 3))
    srv.start()
    drv = ipi.IpiDriver(path, [14] * 4, evaluate=_harmonic, timeout=30)
    assert drv.run(2) == 2
    drv.close()
    srv.join(30)

    path = str(tmp_path / "b.sock")
    srv = Server(path, _frames(4, 1), natom_sent=5)
    srv.start()
    drv = ipi.IpiDriver(path, [14] * 4, evaluate=_harmonic, timeout=30)
    with pytest.raises(ipi.IpiProtocolError, match="mismatch in number of atoms"):
        drv.step()
    drv.close()
    srv.join(30)


def test_unexpected_header(tmp_path):
    path = str(tmp_path / "c.sock")
    srv = Server(path, [], bad_header="INIT")
    srv.start()
    drv = ipi.IpiDriver(path, [14], evaluate=_harmonic, timeout=30)
    with pytest.raises(ipi.IpiProtocolError, match="unexpected message"):
        drv.step()
    drv.close()
    srv.join(30)


def test_refuses_models_without_one_global_target():
    g = fa.GFunctions([2], [8.0], eta=[0.1])
    dims = [1, 2, 2]
    ns = fa.Netstat(g, [14], dims, 0, np.zeros((1, fa.nparams(dims))))
    with pytest.raises(fa.FnetError, match="exactly one global target"):
        ipi.IpiDriver("/nonexistent", [14], ns)


@pytest.mark.gpu
def test_water_box_steps_match_direct_evaluation(tmp_path):
    geos, ns, _ = W.config("water", 0.002)
    ctx = fa.Context(0)
    ctx.upload_model(ns)
    b = ctx.batch(geos[:1], fa.FNET_FP64)
    ns = W.with_zscore(ns, b.acsf())
    b.close()
    ctx.upload_model(ns)
    g0 = geos[0]
    rng = np.random.default_rng(1)
    frames = [(g0.coords + 0.02 * rng.normal(size=g0.coords.shape), g0.latvecs.T)
              for _ in range(3)]
    path = str(tmp_path / "w.sock")
    srv = Server(path, frames)
    srv.start()
    drv = ipi.IpiDriver(path, g0.atnum, ns, precision=fa.FNET_FP32, timeout=120)
    assert drv.run(5) == 3
    drv.close()
    srv.join(60)
    assert srv.error is None, srv.error
    for (coords, h), (e, f, vir, nx) in zip(frames, srv.results):
        geo = fa.Geometry(coords, g0.atnum, g0.species, h.T)
        bb = ctx.batch([geo], fa.FNET_FP32)
        Es, _ = bb.energy()
        F = bb.forces()
        bb.close()
        assert abs(e - Es[0, 0]) <= 1e-9 * max(1.0, abs(Es[0, 0]))
        assert np.abs(f - F).max() <= 1e-9 * max(1.0, np.abs(F).max())
        # and the FP64 oracle, within the FP32 accuracy gates
        from test_gpu_serving import check_against_oracle
        check_against_oracle(ns, geo, e, f, fa.FNET_FP32)
    ctx.close()
