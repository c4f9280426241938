"""FENNIX-shaped model wrapper for the ANI path on the GPU.

Mirrors FeNNol's ``FENNIX`` (models/fennix.py:26-688) for the module set of
the ANI path: preprocessing GRAPH (+ GRAPH_FILTER, GRAPH_ANGULAR_EXTENSION,
SPECIES_INDEXER) and modules ANI_AEV + SPECIES_INDEX_NET (or CHEMICAL_NET).  Same constructor
arguments (``cutoff``, ``modules``, ``preprocessing``, ``variables``,
``energy_terms``, ``energy_unit``), the same entry points (``preprocess``,
``__call__``, ``total_energy``, ``energy_and_forces``,
``energy_and_forces_and_virial``, ``get_gradient_function``) and the same
input/output dictionary keys.  ``rng_key`` is replaced by an integer
``seed`` (weights drawn with NumPy PCG64, see :mod:`fennol_amd.params`), and
``precision`` selects the FP32 or BF16-MLP kernels.  Other FeNNol modules
raise NotImplementedError.
"""

from __future__ import annotations

import re
from collections import OrderedDict

import numpy as np
import torch

from . import units
from .params import (from_chemical_variables, from_variables, init_species_params, to_chemical_variables,
                     to_variables)
from .potential import ANIConfig, ANIPotential
from .preprocessing import (PREPROCESSING, GraphAngularExtension, GraphFilter, GraphGenerator,
                            PreprocessingChain, SpeciesIndexer, atom_unpadding, check_input)

MODULES = {"ANI_AEV": "ANI_AEV", "SPECIES_INDEX_NET": "SPECIES_INDEX_NET", "CHEMICAL_NET": "CHEMICAL_NET"}


def _species_list(v):
    if isinstance(v, str):
        return [el.strip() for el in v.split(",")]
    return list(v)


def _celu_alpha(activation) -> float:
    """Parse the activation string of SPECIES_INDEX_NET (activations.py:296-336); CELU only."""
    if not isinstance(activation, str):
        raise NotImplementedError("only CELU activations are implemented on the GPU path")
    a = activation.replace(" ", "").lower()
    if a == "celu":
        return 1.0
    m = re.fullmatch(r"partial\((?:jax\.nn\.)?celu,alpha=([0-9.eE+-]+)\)", a)
    if m is None:
        raise NotImplementedError(f"activation {activation!r}: only celu / partial(celu,alpha=...) is supported")
    return float(m.group(1))


class FENNIX:
    """ANI FENNIX model evaluated by the HIP kernels.

    Example (the ANI-2x-shaped model of examples/training/ani.yaml:38-82)::

        model = FENNIX(
            cutoff=5.2,
            modules=OrderedDict(
                embedding={"module_name": "ANI_AEV", "species_order": ["H", "C", "N", "O"],
                           "graph_angle_key": "graph_angle", "radial_eta": 19.7, "angular_eta": 12.5,
                           "radial_dist_divisions": 16, "angular_dist_divisions": 4, "zeta": 14.1,
                           "angle_sections": 8},
                energy={"module_name": "SPECIES_INDEX_NET", "output_dim": 1,
                        "hidden_neurons": [256, 192, 160], "species_order": ["H", "C", "N", "O"],
                        "activation": "partial(celu,alpha=0.1)", "input_key": "embedding",
                        "squeeze": True}),
            preprocessing=OrderedDict(
                graph_angle={"module_name": "GRAPH_FILTER", "cutoff": 3.5, "parent_graph": "graph",
                             "graph_key": "graph_angle"},
                angle_extension={"module_name": "GRAPH_ANGULAR_EXTENSION", "graph_key": "graph_angle"},
                species_indexer={"module_name": "SPECIES_INDEXER", "output_key": "species_index"}),
            energy_terms=["energy"], energy_unit="eV", seed=0)
        e, f, out = model.energy_and_forces(species=..., coordinates=...)
    """

    def __init__(self, cutoff: float, modules: OrderedDict, preprocessing: OrderedDict = OrderedDict(),
                 variables: dict | None = None, seed: int | None = 0, energy_terms=None,
                 energy_unit: str = "Ha", precision: str = "fp32", device="cuda", graph_config: dict = {},
                 use_atom_padding: bool = False, **kwargs):
        self._input_args = {"cutoff": cutoff, "modules": OrderedDict(modules),
                            "preprocessing": OrderedDict(preprocessing), "energy_unit": energy_unit}
        self.cutoff = cutoff
        self.energy_unit = energy_unit
        self.Ha_to_model_energy = units.ENERGY_UNITS[energy_unit.upper()]
        self.energy_terms = [energy_terms] if isinstance(energy_terms, str) else energy_terms
        self.precision = precision
        self.use_atom_padding = use_atom_padding

        # ---- preprocessing chain (fennix.py:121-133)
        layers = []
        prep = OrderedDict((k, dict(v)) for k, v in preprocessing.items())
        gparams = {"cutoff": cutoff, "graph_key": "graph"}
        keys = list(prep.keys())
        if keys and keys[0] == "graph":
            gparams.update(prep.pop("graph"))
        gparams.update(graph_config)
        gparams.pop("module_name", None)
        layers.append(GraphGenerator(**{k: v for k, v in gparams.items() if k in ("cutoff", "graph_key",
                                                                                  "mult_size")}))
        for name, params in prep.items():
            key = str(params.pop("module_name", params.pop("FID", name))).upper()
            if key not in PREPROCESSING:
                raise NotImplementedError(f"preprocessing module {key} is not part of the ANI path")
            cls = PREPROCESSING[key]
            if cls is SpeciesIndexer and "species_order" in params and not isinstance(params["species_order"], str):
                params["species_order"] = ",".join(params["species_order"])
            layers.append(cls(**params))

        # ---- modules
        aev_p = sin_p = None
        self.energy_key = None
        for name, params in modules.items():
            params = dict(params)
            key = str(params.pop("module_name", params.pop("FID", name))).upper()
            if key == "ANI_AEV":
                aev_p = (name, params)
            elif key in ("SPECIES_INDEX_NET", "CHEMICAL_NET"):
                if sin_p is not None:
                    raise NotImplementedError("one species network module is supported")
                sin_p = (name, params)
                self.network_kind = key
            else:
                raise NotImplementedError(f"module {key} is not part of the ANI path")
        if aev_p is None or sin_p is None:
            raise ValueError("the ANI path needs an ANI_AEV and a SPECIES_INDEX_NET or CHEMICAL_NET module")
        aname, ap = aev_p
        self.embedding_key = ap.get("embedding_key", "embedding")
        graph_key = ap.get("graph_key", "graph")
        angle_key = ap["graph_angle_key"]
        filt = [l for l in layers if isinstance(l, GraphFilter) and l.graph_key == angle_key]
        if not filt:
            raise ValueError(f"no GRAPH_FILTER produces the angular graph {angle_key!r}")
        if not any(isinstance(l, GraphAngularExtension) and l.graph_key == angle_key for l in layers):
            raise ValueError(f"the angular graph {angle_key!r} needs a GRAPH_ANGULAR_EXTENSION")
        if graph_key != layers[0].graph_key:
            raise NotImplementedError("ANI_AEV must use the main graph")
        self.config = ANIConfig(
            species_order=tuple(_species_list(ap["species_order"])), cutoff=float(cutoff),
            cutoff_angular=float(filt[0].cutoff), radial_eta=float(ap.get("radial_eta", 16.0)),
            angular_eta=float(ap.get("angular_eta", 8.0)),
            radial_dist_divisions=int(ap.get("radial_dist_divisions", 16)),
            angular_dist_divisions=int(ap.get("angular_dist_divisions", 4)), zeta=float(ap.get("zeta", 32.0)),
            angle_sections=int(ap.get("angle_sections", 4)), radial_start=float(ap.get("radial_start", 0.8)),
            angular_start=float(ap.get("angular_start", 0.8)))
        sname, sp = sin_p
        self.energy_module = sname
        self.energy_key = sp.get("output_key") or sname
        chemical = self.network_kind == "CHEMICAL_NET"
        if chemical:
            # ChemicalNet (nets.py:429-512): `neurons` includes the output width; every listed
            # species has a network, atoms of other elements get 0
            neurons = list(sp["neurons"])
            net_species = _species_list(sp["species_order"])
            widths, out_dim = neurons[:-1], neurons[-1]
        else:
            hidden = sp["hidden_neurons"]
            net_species = list(hidden.keys()) if isinstance(hidden, dict) else _species_list(sp["species_order"])
            widths = list(next(iter(hidden.values()))) if isinstance(hidden, dict) else list(hidden)
            out_dim = int(sp.get("output_dim", 1))
        if widths != [256, 192, 160] or int(out_dim) != 1:
            raise NotImplementedError("the MLP kernel implements hidden widths [256, 192, 160] -> 1")
        if not sp.get("use_bias", True):
            raise NotImplementedError("use_bias=False")
        self.celu_alpha = _celu_alpha(sp.get("activation", "silu"))
        if variables is None:
            if seed is None:
                raise ValueError("Either variables or a seed must be provided for initialization")
            params = init_species_params(seed, self.config.aev_length, tuple(widths), 1, tuple(net_species))
            variables = to_chemical_variables(params, sname) if chemical else to_variables(params, sname)
        self.variables = variables
        self.net_species = net_species
        params = from_chemical_variables(variables, net_species, sname) if chemical else from_variables(variables, sname)
        self.potential = ANIPotential(self.config, params, precision=precision, device=device,
                                      unhandled_species="zero" if chemical else "raise")
        if self.celu_alpha != 0.1:
            self.potential.nets.cw.celu_alpha = self.celu_alpha
        self.preprocessing = PreprocessingChain(layers, self.potential, use_atom_padding=use_atom_padding)
        self._graphs_properties = self.preprocessing.get_graphs_properties()
        self.preproc_state = self.preprocessing.init()
        self._cache = {}

    # ------------------------------------------------------------------ state
    def _buffers(self, inputs):
        sp = np.asarray(inputs["species"], np.int32)
        nat = np.asarray(inputs["natoms"], np.int32)
        cells = inputs.get("cells")
        # periodic: the minimum-image convention with flags={"minimum_image": None}, else every
        # image within the cutoff (GraphGenerator, preprocessing.py:127-257)
        minimage = "minimum_image" in inputs.get("flags", {})
        key = (sp.tobytes(), nat.tobytes(), None if cells is None else np.asarray(cells, np.float64).tobytes(),
               minimage)
        sb = self._cache.get(key)
        if sb is None:
            if len(self._cache) > 8:
                self._cache.clear()
            sb = self.potential.setup(sp, nat, cells, minimum_image=minimage)
            self._cache[key] = sb
        return sb

    def preprocess(self, use_gpu: bool = True, **inputs) -> dict:
        """check_input + neighbour lists.  With use_gpu=False the reference's host-path arrays
        (NumPy, padded to capacities) are returned, else the device rows (preprocessing.py:416-442)."""
        inputs = check_input(inputs)
        sb = self._buffers(inputs)
        x32 = torch.as_tensor(np.asarray(inputs["coordinates"], np.float32), device=self.potential.device)
        out = self.preprocessing.process(self.preproc_state, {**inputs, "coordinates": x32}, sb)
        self.preprocessing.check_reallocate(self.preproc_state, out, sb)
        if not use_gpu:
            res = self.preprocessing.export(self.preproc_state, inputs, sb, x32)
            if self.preprocessing.indexer is not None:
                _, res = self.preprocessing.indexer({}, res)
            return res
        return {**out, "coordinates": x32, "_buffers": sb}

    # -------------------------------------------------------------- compute
    def _run(self, unit, forces: bool, embedding: bool = False, **inputs):
        """Energies (and forces, virial) on the device; with use_atom_padding the inputs are padded
        by the chain's AtomPadding first and the outputs unpadded (fennix.py:416-442, 471-540)."""
        inputs = check_input(inputs)
        if self.use_atom_padding:
            self.preproc_state, inputs = self.preprocessing.atom_padding(self.preproc_state, inputs)
        sb = self._buffers(inputs)
        x32 = torch.as_tensor(np.asarray(inputs["coordinates"], np.float32),
                              device=self.potential.device).contiguous()
        self.potential.compute(sb, x32, forces=forces)
        mult = 1.0 if unit is None else (units.ENERGY_UNITS[unit.upper()] if isinstance(unit, str) else unit) \
            / self.Ha_to_model_energy
        out = {**inputs}
        e = sb.energy.cpu().numpy() * mult
        out[self.energy_key] = sb.e_atom.cpu().numpy().astype(np.float64) * mult
        out["atomic_energies"] = np.where(np.asarray(inputs["species"]) > 0, out[self.energy_key], 0.0)
        out["total_energy"] = e
        f = v = None
        if forces:
            f = self.potential.forces(sb).cpu().numpy() * mult
            v = sb.virial.cpu().numpy().reshape(-1, 3, 3) * mult
            out["forces"] = f
            out["virial_tensor"] = v
        if embedding:
            out[self.embedding_key] = sb.aev.float().cpu().numpy()[:, : self.config.aev_length]
        if self.use_atom_padding:
            out = atom_unpadding(out)
            e, f, v = out["total_energy"], out.get("forces"), out.get("virial_tensor")
        return e, f, v, out

    def __call__(self, variables=None, gpu_preprocessing=True, **inputs) -> dict:
        """Model outputs incl. the AEV (``embedding`` key) and atomic energies (fennix.py:466-478)."""
        return self._run(None, False, embedding=True, **inputs)[3]

    def total_energy(self, variables=None, unit=None, gpu_preprocessing=True, **inputs):
        e, _, _, out = self._run(unit, False, **inputs)
        return e, out

    def energy_and_forces(self, variables=None, unit=None, gpu_preprocessing=True, **inputs):
        e, f, _, out = self._run(unit, True, **inputs)
        return e, f, out

    def energy_and_forces_and_virial(self, variables=None, unit=None, gpu_preprocessing=True, **inputs):
        e, f, v, out = self._run(unit, True, **inputs)
        return e, f, v, out

    def get_gradient_function(self, *gradient_keys, jit: bool = True, variables_as_input: bool = False):
        """E and dE/d{coordinates, strain} (fennix.py:365-414)."""
        for k in gradient_keys:
            if k not in ("coordinates", "strain"):
                raise NotImplementedError(f"gradient with respect to {k}")

        def energy_gradient(*args):
            data = args[-1]
            e, f, v, out = self._run(None, True, **data)
            de = {}
            if "coordinates" in gradient_keys:
                de["coordinates"] = -f
            if "strain" in gradient_keys:
                de["strain"] = v
            return e, de, {**out, **{"dEd_" + k: de[k] for k in gradient_keys}}

        return energy_gradient

    @property
    def params(self):
        if self.network_kind == "CHEMICAL_NET":
            return from_chemical_variables(self.variables, self.net_species, self.energy_module)
        return from_variables(self.variables, self.energy_module)

    # ------------------------------------------------------------- files
    def save(self, filename) -> None:
        """Write the model as a .fnx file (msgpack, fennix.py:647-662)."""
        from .fnx import save

        save(self, filename)

    @classmethod
    def load(cls, filename, precision: str = "fp32", device="cuda", graph_config: dict = {},
             use_atom_padding: bool = False) -> "FENNIX":
        """Build a model from a .fnx file (fennix.py:664-688)."""
        from .fnx import load_state

        state = load_state(filename)
        return cls(**state, graph_config=graph_config, use_atom_padding=use_atom_padding, precision=precision,
                   device=device, seed=None)
