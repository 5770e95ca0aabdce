"""vqchip: host-side mirror of VQC.jl's state / circuit / Hamiltonian API over
libvqc_hip.so, the MI355X (gfx950) kernels behind include/vqc_hip.h."""
from . import _lib
from ._lib import VQCError, load
from .api import apply_, apply_operator, circuit_vjp, energy_and_gradient, expectation, expectation_vjp, qubit_encoding_vjp
from .measure import measure, measure_, post_select, post_select_, probability_zero
from .circuit import Program, QCircuit
from .gates import (CONTROL, CONTROLCONTROLGate, CONTROLGate, R, AmplitudeDamping, CCPHASEGate, CNOTGate, CPHASEGate,
                    CRxGate, CRyGate, CRzGate, CZGate, Depolarizing, FREDKINGate, FSIMGate, Gate, HGate, PHASEGate,
                    QuantumGate, QuantumMap, RxGate, RyGate, RzGate, SGate, SWAPGate, TGate, TOFFOLIGate, XGate, YGate,
                    ZGate, from_external, iSWAPdagGate, iSWAPGate)
from .operators import QubitsOperator, QubitsTerm, pack_pauli, pack_terms
from .state import DensityMatrix, StateVector, onehot_encoding, qubit_encoding, rand_state
from .utility import QFT
