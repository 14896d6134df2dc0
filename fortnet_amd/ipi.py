"""i-PI socket driver: the reference's socket-driven mode on the device path.

Fortnet's driver mode (``prg_fnet/fortnet.F90:126-169``) connects to an i-PI server as a client
and, for every geometry step, receives a structure (``ipisocket.F90:175-268`` via
``fnetdata.F90:138-313``), predicts its single global target and the analytic forces
(``predictForSocketComm``, ``fortnet.F90:503-609``) and sends energy, forces and virial back
(``ipisocket.F90:275-359``).  :class:`IpiDriver` speaks the same protocol: 12-byte ASCII headers,
native-endian 32-bit integers and FP64 arrays, all in atomic units.

* ``STATUS`` is answered ``READY`` until ``POSDATA`` (cell as 9 values row after row, its
  inverse, the atom count, the positions) or ``EXIT``; anything else is an error, as is an atom
  count other than the one given at start.
* The result goes out after ``STATUS``/``HAVEDATA`` until ``GETFORCE``: ``FORCEREADY``, the
  energy (the sum of the atomic outputs), the atom count, the forces, the virial and an empty
  extra string.  The reference sends a virial of ones times the cell volume (its stress is set to
  1 with the remark that stress is not supported, ``fortnet.F90:131-132,243-244``); it is sent
  here unchanged so that a run gives the engine the same numbers.
* The model stays on the device and the structure's batch stays resident: after the first step
  only the coordinates and the lattice are uploaded (``fnet_batch_update_coords``), and the
  neighbour search reuses its row capacity.

Models with other than exactly one global target are refused, as in the reference
(``fortnet.F90:541-543``).
"""
import socket
import struct

import numpy as np

from ._lib import FNET_FP32, FnetError
from .core import Context, Geometry

MSGLEN = 12


def _header(text):
    return text.encode("ascii").ljust(MSGLEN)


class IpiProtocolError(RuntimeError):
    pass


class IpiDriver:
    """Client of an i-PI server driving one structure.

    ``address`` is ``(host, port)`` for an internet socket or a path for a UNIX socket (the
    reference takes a port < 1 as UNIX).  ``atomic_numbers`` gives the element of every atom in
    the order the server sends them; ``periodic`` whether the cell is used.  ``evaluate`` may
    replace the device evaluation (a function ``Geometry -> (energy, forces[n, 3])``), which the
    protocol tests use; by default the netstat runs on ``device`` in ``precision``."""

    def __init__(self, address, atomic_numbers, netstat=None, periodic=True,
                 precision=FNET_FP32, device=0, evaluate=None, timeout=None):
        self.atnum = np.asarray(atomic_numbers, np.int32)
        self.n = len(self.atnum)
        if self.n < 1:
            raise ValueError("no atoms")
        self.periodic = bool(periodic)
        self.precision = precision
        self._ctx = None
        self._batch = None
        if evaluate is None:
            if netstat is None:
                raise ValueError("a netstat or an evaluate function is needed")
            if netstat.n_global != 1 or netstat.n_out != 1:
                raise FnetError("socket driven operation only works for exactly one global "
                                "target")
            zs = [int(z) for z in netstat.species_z]
            missing = sorted(set(int(z) for z in self.atnum) - set(zs))
            if missing:
                raise FnetError("the networks have no sub-network for atomic number %d"
                                % missing[0])
            self.species = np.array([zs.index(int(z)) for z in self.atnum], np.int32)
            self._ctx = Context(device)
            self._ctx.upload_model(netstat)
            evaluate = self._device_evaluate
        else:
            self.species = np.zeros(self.n, np.int32)
        self.evaluate = evaluate
        if isinstance(address, (tuple, list)):
            self.sock = socket.create_connection((address[0], int(address[1])), timeout=timeout)
        else:
            self.sock = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
            self.sock.settimeout(timeout)
            self.sock.connect(address)
        self.steps = 0

    # ---- wire ----
    def _recv(self, n):
        buf = bytearray()
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise IpiProtocolError("connection closed by the server")
            buf += chunk
        return bytes(buf)

    def _recv_header(self):
        return self._recv(MSGLEN).decode("ascii").strip()

    def _send_header(self, text):
        self.sock.sendall(_header(text))

    # ---- protocol ----
    def receive(self):
        """Waits for the next structure; returns (coords [n, 3], cell [3, 3] with the lattice
        vectors as rows) or None on EXIT."""
        while True:
            head = self._recv_header()
            if head != "STATUS":
                break
            self._send_header("READY")
        if head == "EXIT":
            return None
        if head != "POSDATA":
            raise IpiProtocolError("unexpected message from server, received '%s'" % head)
        cell = np.frombuffer(self._recv(72), dtype=np.float64).reshape(3, 3)
        self._recv(72)  # inverse cell, not used
        (natom,) = struct.unpack("i", self._recv(4))
        if natom != self.n:
            raise IpiProtocolError("mismatch in number of atoms received: %d, expected %d"
                                   % (natom, self.n))
        coords = np.frombuffer(self._recv(24 * natom), dtype=np.float64).reshape(natom, 3)
        # i-PI sends h row-major with the lattice vectors as columns (h[r][c] = component r of
        # vector c); the reference transposes the 9 values into its latVecs(3, 3) columns
        # (ipisocket.F90:241-244), so the vectors are the rows of this array's transpose
        return coords.copy(), cell.T.copy()

    def send(self, energy, forces, virial):
        """Waits for GETFORCE (answering STATUS with HAVEDATA) and sends the result."""
        while True:
            head = self._recv_header()
            if head != "STATUS":
                break
            self._send_header("HAVEDATA")
        if head != "GETFORCE":
            raise IpiProtocolError("error in socket communication (expected 'GETFORCE', "
                                   "received '%s')" % head)
        forces = np.ascontiguousarray(forces, dtype=np.float64).reshape(self.n, 3)
        msg = [_header("FORCEREADY"), struct.pack("d", float(energy)), struct.pack("i", self.n),
               forces.tobytes(), np.ascontiguousarray(virial, np.float64).reshape(9).tobytes(),
               struct.pack("i", 0)]
        self.sock.sendall(b"".join(msg))

    # ---- evaluation ----
    def _device_evaluate(self, geo):
        if self._batch is None:
            self._batch = self._ctx.batch([geo], self.precision)
        else:
            self._batch.update_coords(geo.coords, geo.latvecs[None] if geo.periodic else None)
        Es, _ = self._batch.energy()
        F = self._batch.forces()
        return float(Es[0, 0]), F.reshape(self.n, 3)

    def step(self):
        """One geometry step; False when the server sent EXIT."""
        got = self.receive()
        if got is None:
            return False
        coords, lat = got
        geo = Geometry(coords, self.atnum, self.species, lat if self.periodic else None)
        energy, forces = self.evaluate(geo)
        vol = abs(float(np.linalg.det(lat))) if self.periodic else 0.0
        self.send(energy, forces, np.full((3, 3), vol))
        self.steps += 1
        return True

    def run(self, n_steps):
        """Serves up to ``n_steps`` geometry steps (the reference's ``nGeoSteps``) or until EXIT;
        returns the number of steps served."""
        for _ in range(n_steps):
            if not self.step():
                break
        return self.steps

    def close(self):
        try:
            self.sock.close()
        finally:
            if self._batch is not None:
                self._batch.close()
                self._batch = None
            if self._ctx is not None:
                self._ctx.close()
                self._ctx = None
