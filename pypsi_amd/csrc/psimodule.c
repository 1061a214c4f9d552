/*
 * psimodule.c -- the Python module "PSI" backed by the MI355X voxelizer.
 *
 * Same surface as the reference module for the density path
 * (src/PSI.c:1071-1111): types Grid and Mesh with the same constructor
 * keywords, defaults, members, error types and messages, and
 * voxels(grid, mesh, mode='density', refine_tolerance=1.0, refine_max_lvl=0),
 * which deposits into grid.fields in place and returns None.  All arithmetic
 * happens on the GPU through libpsivox (include/psivox.h); this file only
 * validates arguments, owns the NumPy arrays and moves pointers.
 *
 * Differences: voxels() also takes kernel='auto'|'plain'|'fast'|'fast64' and
 * xv_true_covariance=False; non-contiguous input arrays are read correctly
 * (the reference reads their raw buffer); the gadget2/gevolution loaders
 * reject malformed particle ids and non-cubic particle counts;
 * RStarTree.query() returns its elements in ascending order (the reference:
 * in its tree's order) and VDF() its streams in ascending element order.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "structmember.h"
#include "../../include/psivox.h"

#define NFIELDS 6
static const char* FIELD_NAMES[NFIELDS] = {"m", "x", "v", "xx", "xv", "vv"};

/* ------------------------------------------------------------------ */
/* Grid (src/PSI.c:48-349)                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD PyObject* type;
    PyObject* fields;
    PyObject* bounds;
    PyObject* n;
    PyObject* d;
    int gtype; /* 0 cart, 1 hpring, -1 invalid */
    int cn[3];
    double win[2][3];
    double cd[3];
    PyObject* arr[NFIELDS]; /* arrays created at init; written by voxels() */
} Grid;

static int Grid_init(Grid* self, PyObject* args, PyObject* kwds) {
    PyObject *type_in = NULL, *bounds_in = NULL, *n_in = NULL, *fields_in = NULL;
    static char* kwlist[] = {"type", "n", "bounds", "fields", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "|UOOO", kwlist, &type_in, &n_in, &bounds_in, &fields_in)) return -1;

    if (type_in) {
        const char* ct = PyUnicode_AsUTF8(type_in);
        if (!ct) return -1;
        if (strcmp(ct, "cart") == 0)
            self->gtype = 0;
        else if (strcmp(ct, "hpring") == 0)
            self->gtype = 1;
        else {
            self->gtype = -1;
            PyErr_Format(PyExc_ValueError, "Invalid grid type: %s", ct);
            return -1;
        }
        Py_INCREF(type_in);
        Py_XSETREF(self->type, type_in);
    }
    if (self->gtype == 1) {
        PyErr_SetString(PyExc_ValueError, "Grid type hpring is not yet implemented");
        return -1;
    }
    if (self->gtype != 0) {
        PyErr_SetString(PyExc_ValueError, "Invalid grid type");
        return -1;
    }
    if (n_in && !PyArg_ParseTuple(n_in, "iii", &self->cn[0], &self->cn[1], &self->cn[2])) {
        PyErr_SetString(PyExc_TypeError, "n must be a tuple containing 3 integers");
        return -1;
    }
    Py_XSETREF(self->n, Py_BuildValue("(iii)", self->cn[0], self->cn[1], self->cn[2]));
    if (bounds_in && !PyArg_ParseTuple(bounds_in, "(ddd)(ddd)", &self->win[0][0], &self->win[0][1], &self->win[0][2],
                                       &self->win[1][0], &self->win[1][1], &self->win[1][2])) {
        PyErr_SetString(PyExc_TypeError, "bounds must be a tuple containing two tuples of 3 values each");
        return -1;
    }
    Py_XSETREF(self->bounds, Py_BuildValue("((ddd)(ddd))", self->win[0][0], self->win[0][1], self->win[0][2],
                                           self->win[1][0], self->win[1][1], self->win[1][2]));
    for (int a = 0; a < 3; ++a) self->cd[a] = (self->win[1][a] - self->win[0][a]) / self->cn[a];
    Py_XSETREF(self->d, Py_BuildValue("(ddd)", self->cd[0], self->cd[1], self->cd[2]));

    npy_intp dims[5] = {self->cn[0], self->cn[1], self->cn[2], 3, 3};
    for (int f = 0; f < NFIELDS; ++f) Py_CLEAR(self->arr[f]);

    PyObject* arr = PyArray_ZEROS(3, dims, NPY_FLOAT64, 0);
    if (!arr) return -1;
    self->arr[0] = arr;
    if (PyDict_SetItemString(self->fields, "m", arr) < 0) return -1;

    if (fields_in != NULL && fields_in != Py_None) {
        PyObject* seq = PySequence_Fast(fields_in, "Expected a sequence for fields.");
        if (!seq) return -1;
        Py_ssize_t len = PySequence_Fast_GET_SIZE(seq);
        for (Py_ssize_t i = 0; i < len; ++i) {
            PyObject* it = PySequence_Fast_GET_ITEM(seq, i);
            if (!PyUnicode_Check(it)) {
                Py_DECREF(seq);
                PyErr_SetString(PyExc_TypeError, "Field identifiers must be strings");
                return -1;
            }
            const char* cs = PyUnicode_AsUTF8(it);
            int slot = -1, nd = 0;
            if (strcmp(cs, "m") == 0) continue;
            if (strcmp(cs, "x") == 0) {
                slot = 1;
                nd = 4;
            } else if (strcmp(cs, "v") == 0) {
                slot = 2;
                nd = 4;
            } else if (strcmp(cs, "xx") == 0) {
                slot = 3;
                nd = 5;
            } else if (strcmp(cs, "xv") == 0 || strcmp(cs, "vx") == 0) {
                slot = 4;
                nd = 5;
            } else if (strcmp(cs, "vv") == 0) {
                slot = 5;
                nd = 5;
            } else {
                Py_DECREF(seq);
                PyErr_Format(PyExc_ValueError, "Unrecognized field identifier: %s", cs);
                return -1;
            }
            PyObject* a = PyArray_ZEROS(nd, dims, NPY_FLOAT64, 0);
            if (!a) {
                Py_DECREF(seq);
                return -1;
            }
            Py_XSETREF(self->arr[slot], a);
            if (PyDict_SetItemString(self->fields, FIELD_NAMES[slot], a) < 0) {
                Py_DECREF(seq);
                return -1;
            }
        }
        Py_DECREF(seq);
        if (self->arr[3] && !self->arr[1]) {
            PyErr_SetString(PyExc_ValueError, "Field 'xx' requires you to also provide field 'x'.");
            return -1;
        }
        if (self->arr[4] && !(self->arr[2] && self->arr[1])) {
            PyErr_SetString(PyExc_ValueError, "Field 'xv' requires you to also provide fields 'x' and 'v'.");
            return -1;
        }
        if (self->arr[5] && !self->arr[2]) {
            PyErr_SetString(PyExc_ValueError, "Field 'vv' requires you to also provide field 'v'.");
            return -1;
        }
    }
    return 0;
}

static PyObject* Grid_new(PyTypeObject* type, PyObject* args, PyObject* kwds) {
    Grid* self = (Grid*)type->tp_alloc(type, 0);
    if (!self) return NULL;
    self->gtype = 0;
    for (int a = 0; a < 3; ++a) {
        self->cn[a] = 32;
        self->win[0][a] = 0.0;
        self->win[1][a] = 1.0;
        self->cd[a] = (self->win[1][a] - self->win[0][a]) / self->cn[a];
    }
    self->fields = PyDict_New();
    self->type = PyUnicode_FromString("cart");
    Py_INCREF(Py_None);
    self->bounds = Py_None;
    Py_INCREF(Py_None);
    self->n = Py_None;
    Py_INCREF(Py_None);
    self->d = Py_None;
    if (!self->fields || !self->type) {
        Py_DECREF(self);
        return NULL;
    }
    return (PyObject*)self;
}

static void Grid_dealloc(Grid* self) {
    Py_XDECREF(self->type);
    Py_XDECREF(self->fields);
    Py_XDECREF(self->bounds);
    Py_XDECREF(self->n);
    Py_XDECREF(self->d);
    for (int f = 0; f < NFIELDS; ++f) Py_XDECREF(self->arr[f]);
    Py_TYPE(self)->tp_free((PyObject*)self);
}

static PyMemberDef Grid_members[] = {
    {"type", T_OBJECT, offsetof(Grid, type), 0, "type"},
    {"fields", T_OBJECT, offsetof(Grid, fields), 0, "fields dict"},
    {"bounds", T_OBJECT, offsetof(Grid, bounds), 0, "lower-left and upper-right corners of the grid"},
    {"n", T_OBJECT, offsetof(Grid, n), 0, "number of gridpoints"},
    {"d", T_OBJECT, offsetof(Grid, d), 0, "grid dx"},
    {NULL}};

static PyTypeObject GridType = {
    PyVarObject_HEAD_INIT(NULL, 0).tp_name = "PSI.Grid",
    .tp_basicsize = sizeof(Grid),
    .tp_dealloc = (destructor)Grid_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT | Py_TPFLAGS_BASETYPE,
    .tp_doc = "Basic grid",
    .tp_members = Grid_members,
    .tp_init = (initproc)Grid_init,
    .tp_new = Grid_new,
};

/* ------------------------------------------------------------------ */
/* Mesh (src/PSI.c:358-806)                                            */
/* ------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD PyObject* pos;
    PyObject* vel;
    PyObject* mass;
    PyObject* connectivity;
    PyObject* wrapbox;
    int elemtype;
    int periodic;
    int64_t npart, nelem;
    double box[2][3];
} Mesh;

static int is_f64(PyObject* o) { return PyArray_Check(o) && PyArray_TYPE((PyArrayObject*)o) == NPY_FLOAT64; }

static PyObject* zeros_like_shape(int nd, npy_intp* dims) { return PyArray_ZEROS(nd, dims, NPY_FLOAT64, 0); }

/* trilinear lattice connectivity: element (i, j, k) of an ne lattice (C
 * order) takes particle ((i+ii) mod ns0, (j+jj) mod ns1, (k+kk) mod ns2) of
 * the ns lattice as corner 4ii+2jj+kk (src/PSI.c:636-656, src/mesh.c:137-150) */
static void lattice_connectivity(int32_t* cd, const npy_intp* ns, const npy_intp* ne) {
    int32_t* row = cd;
    for (npy_intp i = 0; i < ne[0]; ++i) {
        const npy_intp ia[2] = {i * ns[1], (i + 1 == ns[0] ? 0 : i + 1) * ns[1]};
        for (npy_intp j = 0; j < ne[1]; ++j) {
            const npy_intp jb[2] = {j, j + 1 == ns[1] ? 0 : j + 1};
            for (npy_intp k = 0; k < ne[2]; ++k, row += 8) {
                const npy_intp kc[2] = {k, k + 1 == ns[2] ? 0 : k + 1};
                for (int c = 0; c < 8; ++c)
                    row[c] = (int32_t)((ia[c >> 2] + jb[(c >> 1) & 1]) * ns[2] + kc[c & 1]);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Snapshot loaders: Gadget-2 format-1 files, particle type 1          */
/* (src/mesh.c:35-302, src/PSI.c:659-707)                              */
/* ------------------------------------------------------------------ */

/* The header block is 256 bytes; the loaders use npart[1] (int32 at byte
 * 4), mass[1] (double at byte 32) and BoxSize (double at byte 128). */
enum { HDR_BYTES = 256, HDR_NPART1 = 4, HDR_MASS1 = 32, HDR_BOX = 128 };

/* one Fortran record holding exactly nbytes: both 4-byte length words must
 * say nbytes (the block-size checks of src/mesh.c:101-121) */
static int read_record(FILE* f, void* dst, size_t nbytes) {
    uint32_t head, tail;
    if (fread(&head, 4, 1, f) != 1 || (size_t)head != nbytes) return 0;
    if (nbytes && fread(dst, 1, nbytes, f) != nbytes) return 0;
    return fread(&tail, 4, 1, f) == 1 && (size_t)tail == nbytes;
}

/* gevolution writes particles in tiles of 4x4x4, tile (ti, tj, tk) at index
 * ti + T tj + T^2 tk.  Inside a tile, particle f sits at 2 G(f / 8) + G(f % 8)
 * with G(n) = n ^ (n >> 1), the 3-bit reflected Gray code, read as
 * (i, j, k) = (bit 2, bit 1, bit 0): the order of the tables at
 * src/mesh.c:163-164. */
static void tile_position(int f, int* c) {
    const int go = (f >> 3) ^ (f >> 4), gi = (f & 7) ^ ((f & 7) >> 1);
    for (int a = 0; a < 3; ++a) c[a] = 2 * ((go >> (2 - a)) & 1) + ((gi >> (2 - a)) & 1);
}

/* element p of a gevolution mesh is the lattice cube whose corner 0 is
 * particle p; corner 4ii+2jj+kk is the particle (ii, jj, kk) further on,
 * across tile edges and the periodic box (src/mesh.c:168-209, 296-299) */
static void tiled_connectivity(int32_t* cd, int64_t T) {
    /* the in-tile cell of each slot and the slot of each cell, once */
    int pos64[64][3], slot64[4][4][4];
    for (int f = 0; f < 64; ++f) {
        tile_position(f, pos64[f]);
        slot64[pos64[f][0]][pos64[f][1]][pos64[f][2]] = f;
    }
    int32_t* row = cd;
    for (int64_t tk = 0; tk < T; ++tk)
        for (int64_t tj = 0; tj < T; ++tj)
            for (int64_t ti = 0; ti < T; ++ti) {
                /* per axis: this tile's coordinate and its +1 neighbour's */
                const int64_t t[3] = {ti, tj, tk};
                int64_t base[3][2];
                for (int a = 0; a < 3; ++a) {
                    base[a][0] = t[a];
                    base[a][1] = t[a] + 1 == T ? 0 : t[a] + 1;
                }
                for (int f = 0; f < 64; ++f, row += 8) {
                    const int* c = pos64[f];
                    for (int corner = 0; corner < 8; ++corner) {
                        const int s0 = c[0] + (corner >> 2), s1 = c[1] + ((corner >> 1) & 1), s2 = c[2] + (corner & 1);
                        const int64_t u0 = base[0][s0 >> 2], u1 = base[1][s1 >> 2], u2 = base[2][s2 >> 2];
                        row[corner] = (int32_t)((u0 + T * u1 + T * T * u2) * 64 + slot64[s0 & 3][s1 & 3][s2 & 3]);
                    }
                }
            }
}

/* loader='gadget2' or 'gevolution': positions and velocities (float32),
 * particle ids (int32 for gadget2, int64 for gevolution), no mass block.
 * Particle id i goes to row i; every element gets mass[1] / npart; the
 * mesh is periodic on [0, BoxSize]^3 (src/mesh.c:62-155, 214-302).
 * Errors as src/PSI.c:665-702: ValueError "Bad file name" when the file
 * does not open, "Could not read" when its blocks are not as described.
 * Unlike the reference, ids that are not a permutation of 0..npart-1 and a
 * particle count that is not a cube (of a multiple of 4 for gevolution's
 * tiles) are errors too; the reference writes out of bounds or leaves
 * connectivity rows unset there. */
static int load_snapshot(Mesh* self, const char* loader, const char* fn) {
    const int gev = strcmp(loader, "gevolution") == 0;
    PySys_WriteStdout(gev ? "Using the Gevolution loader.\n" : "Using the Gadget2 loader.\n");
    if (!fn) {
        PyErr_SetString(PyExc_ValueError, "Bad file name: None");
        return -1;
    }
    FILE* f = fopen(fn, "rb");
    if (!f) {
        PySys_WriteStdout("Failed to open %.900s.\n", fn);
        PyErr_Format(PyExc_ValueError, "Bad file name: %s", fn);
        return -1;
    }
    PySys_WriteStdout("Loading %.900s...\n", fn);
    const char* why = NULL;
    unsigned char hdr[HDR_BYTES];
    uint32_t mark;
    int32_t np = 0;
    double m1 = 0.0, boxsz = 0.0;
    if (fread(&mark, 4, 1, f) != 1 || fread(hdr, 1, HDR_BYTES, f) != HDR_BYTES || fread(&mark, 4, 1, f) != 1)
        why = "short header";
    else {
        memcpy(&np, hdr + HDR_NPART1, 4);
        memcpy(&m1, hdr + HDR_MASS1, 8);
        memcpy(&boxsz, hdr + HDR_BOX, 8);
        if (np < 0) why = "negative particle count";
    }
    const int64_t n = np;
    int64_t side = (int64_t)floor(cbrt((double)n) + 0.5);
    if (!why && (side * side * side != n || (gev && side % 4 != 0)))
        why = gev ? "particle count is not a cube of a multiple of 4" : "particle count is not a cube";
    const size_t idb = gev ? 8 : 4;
    float *tp = NULL, *tv = NULL;
    void* tid = NULL;
    if (!why) {
        tp = (float*)malloc(12 * (size_t)n + 4);
        tv = (float*)malloc(12 * (size_t)n + 4);
        tid = malloc(idb * (size_t)n + 8);
        if (!tp || !tv || !tid) why = "out of host memory";
    }
    if (!why) {
        int ok;
        Py_BEGIN_ALLOW_THREADS;
        ok = read_record(f, tp, 12 * (size_t)n) && read_record(f, tv, 12 * (size_t)n) &&
             read_record(f, tid, idb * (size_t)n);
        Py_END_ALLOW_THREADS;
        if (!ok) why = "block sizes do not match the header";
    }
    fclose(f);
    PyObject *pos = NULL, *vel = NULL, *mass = NULL, *conn = NULL;
    if (!why) {
        npy_intp d2[2] = {(npy_intp)n, 3}, dc[2] = {(npy_intp)n, 8};
        pos = PyArray_SimpleNew(2, d2, NPY_FLOAT64);
        vel = PyArray_SimpleNew(2, d2, NPY_FLOAT64);
        mass = PyArray_SimpleNew(1, d2, NPY_FLOAT64);
        conn = PyArray_SimpleNew(2, dc, NPY_INT32);
        if (!pos || !vel || !mass || !conn) why = "out of host memory";
    }
    if (!why) {
        double* pd = (double*)PyArray_DATA((PyArrayObject*)pos);
        double* vd = (double*)PyArray_DATA((PyArrayObject*)vel);
        double* md = (double*)PyArray_DATA((PyArrayObject*)mass);
        int32_t* cd = (int32_t*)PyArray_DATA((PyArrayObject*)conn);
        unsigned char* seen = (unsigned char*)calloc((size_t)n + 1, 1);
        if (!seen) why = "out of host memory";
        Py_BEGIN_ALLOW_THREADS;
        for (int64_t p = 0; !why && p < n; ++p) {
            const int64_t id = gev ? ((const int64_t*)tid)[p] : (int64_t)((const int32_t*)tid)[p];
            if (id < 0 || id >= n || seen[id]) {
                why = "particle ids are not a permutation of 0..npart-1";
                break;
            }
            seen[id] = 1;
            for (int a = 0; a < 3; ++a) {
                pd[3 * id + a] = (double)tp[3 * p + a];
                vd[3 * id + a] = (double)tv[3 * p + a];
            }
            md[id] = m1 / (double)n;
        }
        if (!why) {
            if (gev)
                tiled_connectivity(cd, side / 4);
            else {
                const npy_intp ns[3] = {(npy_intp)side, (npy_intp)side, (npy_intp)side};
                lattice_connectivity(cd, ns, ns);
            }
        }
        Py_END_ALLOW_THREADS;
        free(seen);
    }
    free(tp);
    free(tv);
    free(tid);
    if (why) {
        Py_XDECREF(pos);
        Py_XDECREF(vel);
        Py_XDECREF(mass);
        Py_XDECREF(conn);
        if (!PyErr_Occurred()) {
            PySys_WriteStdout("%.900s: %s\n", fn, why);
            PyErr_Format(PyExc_ValueError, "Could not read %s", fn);
        }
        return -1;
    }
    Py_XSETREF(self->pos, pos);
    Py_XSETREF(self->vel, vel);
    Py_XSETREF(self->mass, mass);
    Py_XSETREF(self->connectivity, conn);
    self->npart = n;
    self->nelem = n;
    self->elemtype = 8;
    self->periodic = 1;
    for (int a = 0; a < 3; ++a) {
        self->box[0][a] = 0.0;
        self->box[1][a] = boxsz;
    }
    PySys_WriteStdout("...done.\n");
    return 0;
}

static int Mesh_init(Mesh* self, PyObject* args, PyObject* kwds) {
    char *cloader, *cfile = NULL;
    PyObject *posar = NULL, *velar = NULL, *massar = NULL, *connar = NULL, *box = NULL;
    static char* kwlist[] = {"loader", "filename", "pos", "vel", "mass", "connectivity", "periodic_box", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "s|sOOOOO", kwlist, &cloader, &cfile, &posar, &velar, &massar,
                                     &connar, &box))
        return -1;

    if (strcmp(cloader, "array") == 0 || strcmp(cloader, "block") == 0) {
        if (box) {
            if (!PyArg_ParseTuple(box, "(ddd)(ddd)", &self->box[0][0], &self->box[0][1], &self->box[0][2],
                                  &self->box[1][0], &self->box[1][1], &self->box[1][2])) {
                PyErr_SetString(PyExc_TypeError, "periodic_box must be a tuple containing two tuples of 3 values each");
                return -1;
            }
            self->periodic = 1;
        } else {
            self->periodic = 0;
        }
    }

    if (strcmp(cloader, "array") == 0) {
        if (!PyArray_Check(posar)) {
            PyErr_SetString(PyExc_TypeError, "pos must be a numpy array");
            return -1;
        }
        if (!is_f64(posar)) {
            PyErr_SetString(PyExc_TypeError, "pos must be an array of doubles");
            return -1;
        }
        PyArrayObject* pa = (PyArrayObject*)posar;
        if (PyArray_NDIM(pa) != 2 || PyArray_DIMS(pa)[1] != 3) {
            PyErr_SetString(PyExc_TypeError, "pos must be a 2D array with shape (npart, 3)");
            return -1;
        }
        self->npart = PyArray_DIMS(pa)[0];
        Py_INCREF(posar);
        Py_XSETREF(self->pos, posar);
        if (!PyArray_Check(connar)) {
            PyErr_SetString(PyExc_TypeError, "connectivity must be a numpy array");
            return -1;
        }
        PyArrayObject* ca = (PyArrayObject*)connar;
        if (PyArray_TYPE(ca) != NPY_INT32) {
            PyErr_SetString(PyExc_TypeError, "connectivity must be an array of ints");
            return -1;
        }
        if (PyArray_NDIM(ca) != 2) {
            PyErr_SetString(PyExc_TypeError, "connectivity must be a 2D array with shape (nelem, nverts_per_elem)");
            return -1;
        }
        npy_intp vpe = PyArray_DIMS(ca)[1];
        if (vpe != 4 && vpe != 8 && vpe != 27) {
            PyErr_SetString(PyExc_TypeError,
                            "connectivity must be a 2D array where nverts_per_elem is 4, 8, or 27.");
            return -1;
        }
        self->elemtype = (int)vpe;
        self->nelem = PyArray_DIMS(ca)[0];
        {
            /* the reference reads pos[conn] unchecked (src/cpsi.c:85-98); here
             * an entry outside [0, npart) is an error, not a read past the
             * array (the library checks device meshes the same way) */
            PyArrayObject* cc = PyArray_GETCONTIGUOUS(ca);
            if (!cc) return -1;
            const int32_t* cd = (const int32_t*)PyArray_DATA(cc);
            const npy_intp nc = PyArray_SIZE(cc);
            npy_intp bad = -1;
            for (npy_intp i = 0; i < nc && bad < 0; ++i)
                if (cd[i] < 0 || (int64_t)cd[i] >= self->npart) bad = i;
            const long long bv = bad >= 0 ? (long long)cd[bad] : 0;
            Py_DECREF(cc);
            if (bad >= 0) {
                PyErr_Format(PyExc_ValueError, "connectivity entry %lld (element %lld) = %lld is outside [0, %lld)",
                             (long long)bad, (long long)(bad / vpe), bv, (long long)self->npart);
                return -1;
            }
        }
        Py_INCREF(connar);
        Py_XSETREF(self->connectivity, connar);
        if (velar == NULL) {
            npy_intp d2[2] = {(npy_intp)self->npart, 3};
            PyObject* v = zeros_like_shape(2, d2);
            if (!v) return -1;
            Py_XSETREF(self->vel, v);
        } else if (PyArray_Check(velar)) {
            if (!is_f64(velar)) {
                PyErr_SetString(PyExc_TypeError, "if provided, vel must be an array of doubles");
                return -1;
            }
            PyArrayObject* va = (PyArrayObject*)velar;
            if (PyArray_NDIM(va) != 2 || PyArray_DIMS(va)[1] != 3 || PyArray_DIMS(va)[0] != self->npart) {
                PyErr_SetString(PyExc_TypeError, "if provided, vel must have the same shape as pos");
                return -1;
            }
            Py_INCREF(velar);
            Py_XSETREF(self->vel, velar);
        } else {
            PyErr_SetString(PyExc_TypeError, "if provided, vel must be a numpy array");
            return -1;
        }
        if (massar == NULL) {
            npy_intp d1[1] = {(npy_intp)self->nelem};
            PyObject* m = PyArray_SimpleNew(1, d1, NPY_FLOAT64);
            if (!m) return -1;
            double* md = (double*)PyArray_DATA((PyArrayObject*)m);
            for (int64_t i = 0; i < self->nelem; ++i) md[i] = 1.0;
            Py_XSETREF(self->mass, m);
        } else if (PyArray_Check(massar)) {
            if (!is_f64(massar)) {
                PyErr_SetString(PyExc_TypeError, "if provided, mass must be an array of doubles");
                return -1;
            }
            PyArrayObject* ma = (PyArrayObject*)massar;
            if (PyArray_NDIM(ma) != 1 || PyArray_DIMS(ma)[0] != self->nelem) {
                PyErr_SetString(PyExc_TypeError, "if provided, mass must be a 1D array of length nelem");
                return -1;
            }
            Py_INCREF(massar);
            Py_XSETREF(self->mass, massar);
        } else {
            PyErr_SetString(PyExc_TypeError, "if provided, mass must be a numpy array");
            return -1;
        }
    } else if (strcmp(cloader, "block") == 0) {
        if (!PyArray_Check(posar)) {
            PyErr_SetString(PyExc_TypeError, "pos must be a numpy array");
            return -1;
        }
        if (!is_f64(posar)) {
            PyErr_SetString(PyExc_TypeError, "pos must be an array of doubles");
            return -1;
        }
        PyArrayObject* pa = (PyArrayObject*)posar;
        if (PyArray_NDIM(pa) != 4 || PyArray_DIMS(pa)[3] != 3) {
            PyErr_SetString(PyExc_TypeError, "pos must be a 4D array with shape (nx, ny, nz, 3)");
            return -1;
        }
        npy_intp ns[3] = {PyArray_DIMS(pa)[0], PyArray_DIMS(pa)[1], PyArray_DIMS(pa)[2]};
        self->npart = (int64_t)ns[0] * ns[1] * ns[2];
        Py_INCREF(posar);
        Py_XSETREF(self->pos, posar);
        npy_intp ne[3];
        for (int a = 0; a < 3; ++a) ne[a] = self->periodic ? ns[a] : ns[a] - 1;
        self->nelem = (int64_t)ne[0] * ne[1] * ne[2];
        self->elemtype = 8;
        if (velar == NULL) {
            npy_intp d4[4] = {ns[0], ns[1], ns[2], 3};
            PyObject* v = zeros_like_shape(4, d4);
            if (!v) return -1;
            Py_XSETREF(self->vel, v);
        } else if (PyArray_Check(velar)) {
            if (!is_f64(velar)) {
                PyErr_SetString(PyExc_TypeError, "if provided, vel must be an array of doubles");
                return -1;
            }
            PyArrayObject* va = (PyArrayObject*)velar;
            if (PyArray_NDIM(va) != 4 || PyArray_DIMS(va)[3] != 3 || PyArray_DIMS(va)[0] != ns[0] ||
                PyArray_DIMS(va)[1] != ns[1] || PyArray_DIMS(va)[2] != ns[2]) {
                PyErr_SetString(PyExc_TypeError, "if provided, vel must have the same shape as pos");
                return -1;
            }
            Py_INCREF(velar);
            Py_XSETREF(self->vel, velar);
        } else {
            PyErr_SetString(PyExc_TypeError, "if provided, vel must be a numpy array");
            return -1;
        }
        if (massar == NULL) {
            PyObject* m = PyArray_SimpleNew(3, ne, NPY_FLOAT64);
            if (!m) return -1;
            double* md = (double*)PyArray_DATA((PyArrayObject*)m);
            for (int64_t i = 0; i < self->nelem; ++i) md[i] = 1.0;
            Py_XSETREF(self->mass, m);
        } else if (PyArray_Check(massar)) {
            if (!is_f64(massar)) {
                PyErr_SetString(PyExc_TypeError, "if provided, mass must be an array of doubles");
                return -1;
            }
            PyArrayObject* ma = (PyArrayObject*)massar;
            if (PyArray_NDIM(ma) != 3 || PyArray_DIMS(ma)[0] != ne[0] || PyArray_DIMS(ma)[1] != ne[1] ||
                PyArray_DIMS(ma)[2] != ne[2]) {
                PyErr_SetString(PyExc_TypeError,
                                "if provided, mass must be a 3D array of shape (nelemx, nelemy, nelemz)");
                return -1;
            }
            Py_INCREF(massar);
            Py_XSETREF(self->mass, massar);
        } else {
            PyErr_SetString(PyExc_TypeError, "if provided, mass must be a numpy array");
            return -1;
        }
        /* trilinear connectivity, corner order 4ii+2jj+kk, wrapping modulo
         * the lattice when periodic (src/PSI.c:636-656) */
        npy_intp dc[2] = {(npy_intp)self->nelem, 8};
        PyObject* c = PyArray_SimpleNew(2, dc, NPY_INT32);
        if (!c) return -1;
        lattice_connectivity((int32_t*)PyArray_DATA((PyArrayObject*)c), ns, ne);
        Py_XSETREF(self->connectivity, c);
    } else if (strcmp(cloader, "gadget2") == 0 || strcmp(cloader, "gevolution") == 0) {
        if (load_snapshot(self, cloader, cfile) < 0) return -1;
    } else {
        PyErr_SetString(PyExc_ValueError, "Invalid mesh loader.");
        return -1;
    }
    if (self->periodic) {
        Py_XSETREF(self->wrapbox, Py_BuildValue("((ddd)(ddd))", self->box[0][0], self->box[0][1], self->box[0][2],
                                                self->box[1][0], self->box[1][1], self->box[1][2]));
    }
    return 0;
}

static PyObject* Mesh_new(PyTypeObject* type, PyObject* args, PyObject* kwds) {
    Mesh* self = (Mesh*)type->tp_alloc(type, 0);
    if (!self) return NULL;
    PyObject** slots[5] = {&self->pos, &self->vel, &self->mass, &self->connectivity, &self->wrapbox};
    for (int i = 0; i < 5; ++i) {
        Py_INCREF(Py_None);
        *slots[i] = Py_None;
    }
    return (PyObject*)self;
}

static void Mesh_dealloc(Mesh* self) {
    Py_XDECREF(self->pos);
    Py_XDECREF(self->vel);
    Py_XDECREF(self->mass);
    Py_XDECREF(self->connectivity);
    Py_XDECREF(self->wrapbox);
    Py_TYPE(self)->tp_free((PyObject*)self);
}

static PyMemberDef Mesh_members[] = {{"pos", T_OBJECT_EX, offsetof(Mesh, pos), 0, "pos"},
                                     {"vel", T_OBJECT_EX, offsetof(Mesh, vel), 0, "vel"},
                                     {"mass", T_OBJECT_EX, offsetof(Mesh, mass), 0, "mass"},
                                     {"connectivity", T_OBJECT_EX, offsetof(Mesh, connectivity), 0, "connectivity"},
                                     {"wrapbox", T_OBJECT_EX, offsetof(Mesh, wrapbox), 0, "wrapbox"},
                                     {NULL}};

static PyTypeObject MeshType = {
    PyVarObject_HEAD_INIT(NULL, 0).tp_name = "PSI.Mesh",
    .tp_basicsize = sizeof(Mesh),
    .tp_dealloc = (destructor)Mesh_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT | Py_TPFLAGS_BASETYPE,
    .tp_doc = "Mesh",
    .tp_members = Mesh_members,
    .tp_init = (initproc)Mesh_init,
    .tp_new = Mesh_new,
};

/* ------------------------------------------------------------------ */
/* voxels (src/PSI.c:974-1018)                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t last_print;
    int verbose;
} progress_state;

/* between device batches: progress line (src/cpsi.c:82-83) and Ctrl-C
 * (src/cpsi.c:160-163) */
static int progress_cb(void* ctx, int64_t done, int64_t total) {
    progress_state* ps = (progress_state*)ctx;
    PyGILState_STATE gs = PyGILState_Ensure();
    if (ps->verbose && done < total)
        PySys_WriteStdout("\rElement %lld of %lld, %.1f%%", (long long)done, (long long)total,
                          total ? (100.0 * done) / total : 100.0);
    int stop = PyErr_CheckSignals() < 0;
    PyGILState_Release(gs);
    return stop;
}

static PyObject* contiguous(PyObject* o) {
    return (PyObject*)PyArray_GETCONTIGUOUS((PyArrayObject*)o);
}

static PyObject* PSI_voxels(PyObject* self, PyObject* args, PyObject* kwds) {
    (void)self;
    Grid* grid;
    Mesh* mesh;
    const char* modestr = "density";
    const char* kernstr = "auto";
    double reftol = 1.0;
    int maxlvl = 0, xvfix = 0;
    static char* kwlist[] = {"grid", "mesh", "mode", "refine_tolerance", "refine_max_lvl", "kernel",
                             "xv_true_covariance", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!O!|sdi$sp", kwlist, &GridType, &grid, &MeshType, &mesh, &modestr,
                                     &reftol, &maxlvl, &kernstr, &xvfix))
        return NULL;
    if (grid->gtype != 0) {
        PyErr_SetString(PyExc_ValueError, "Only cartesian grids are supported by PSI.voxels()");
        return NULL;
    }
    if (mesh->pos == Py_None) {
        PyErr_SetString(PyExc_ValueError, "Input mesh and grid must have the same number of dimensions");
        return NULL;
    }
    int mode;
    if (strcmp(modestr, "density") == 0)
        mode = PSIVOX_MODE_DENSITY;
    else if (strcmp(modestr, "annihilation") == 0)
        mode = PSIVOX_MODE_ANNIHILATION;
    else {
        PyErr_Format(PyExc_ValueError, "Invalid mode for PSI.voxels(): %s", modestr);
        return NULL;
    }
    int kernel;
    if (strcmp(kernstr, "auto") == 0)
        kernel = PSIVOX_KERNEL_AUTO;
    else if (strcmp(kernstr, "plain") == 0)
        kernel = PSIVOX_KERNEL_PLAIN;
    else if (strcmp(kernstr, "fast") == 0)
        kernel = PSIVOX_KERNEL_FAST;
    else if (strcmp(kernstr, "fast64") == 0)
        kernel = PSIVOX_KERNEL_FAST64;
    else {
        PyErr_Format(PyExc_ValueError, "Invalid kernel for PSI.voxels(): %s", kernstr);
        return NULL;
    }
    if (psivox_device_count() <= 0) {
        PyErr_SetString(PyExc_RuntimeError, "PSI.voxels(): no GPU visible to the MI355X voxelizer");
        return NULL;
    }

    PyObject *pos = NULL, *vel = NULL, *mass = NULL, *conn = NULL;
    PyObject* fa[NFIELDS] = {NULL};
    PyObject* ret = NULL;
    pos = contiguous(mesh->pos);
    vel = contiguous(mesh->vel);
    mass = contiguous(mesh->mass);
    conn = contiguous(mesh->connectivity);
    if (!pos || !vel || !mass || !conn) goto done;

    psivox_mesh cm;
    memset(&cm, 0, sizeof cm);
    cm.elemtype = mesh->elemtype;
    cm.periodic = mesh->periodic;
    cm.npart = mesh->npart;
    cm.nelem = mesh->nelem;
    cm.pos = (const double*)PyArray_DATA((PyArrayObject*)pos);
    cm.vel = (const double*)PyArray_DATA((PyArrayObject*)vel);
    cm.mass = (const double*)PyArray_DATA((PyArrayObject*)mass);
    cm.conn = (const int32_t*)PyArray_DATA((PyArrayObject*)conn);
    memcpy(cm.box, mesh->box, sizeof cm.box);

    psivox_grid cg;
    memset(&cg, 0, sizeof cg);
    for (int a = 0; a < 3; ++a) {
        cg.n[a] = grid->cn[a];
        cg.window[0][a] = grid->win[0][a];
        cg.window[1][a] = grid->win[1][a];
        cg.d[a] = grid->cd[a];
    }
    for (int f = 0; f < NFIELDS; ++f) {
        if (!grid->arr[f]) continue;
        fa[f] = grid->arr[f];
        Py_INCREF(fa[f]);
        cg.fields[f] = (double*)PyArray_DATA((PyArrayObject*)fa[f]);
    }

    psivox_opts opts;
    memset(&opts, 0, sizeof opts);
    opts.mode = mode;
    opts.refine_tolerance = reftol;
    opts.refine_max_lvl = maxlvl;
    opts.kernel = kernel;
    opts.xv_true_covariance = xvfix;
    progress_state ps = {0, 1};
    opts.progress = progress_cb;
    opts.progress_ctx = &ps;

    psivox_stats st;
    int rc;
    Py_BEGIN_ALLOW_THREADS rc = psivox_voxels(&cm, &cg, &opts, &st);
    Py_END_ALLOW_THREADS if (rc == PSIVOX_E_CANCELLED) {
        if (!PyErr_Occurred()) PyErr_SetString(PyExc_KeyboardInterrupt, psivox_last_error());
        goto done;
    }
    if (rc != 0) {
        PyErr_Format(rc == PSIVOX_E_NOMEM ? PyExc_MemoryError : PyExc_RuntimeError, "PSI.voxels(): %s",
                     psivox_last_error());
        goto done;
    }
    PySys_WriteStdout("\rElement %lld of %lld, %.1f%%\n", (long long)mesh->nelem, (long long)mesh->nelem, 100.0);
    Py_INCREF(Py_None);
    ret = Py_None;
done:
    Py_XDECREF(pos);
    Py_XDECREF(vel);
    Py_XDECREF(mass);
    Py_XDECREF(conn);
    for (int f = 0; f < NFIELDS; ++f) Py_XDECREF(fa[f]);
    return ret;
}

/* ------------------------------------------------------------------ */
/* RStarTree (src/PSI.c:808-935): the mesh's element boxes, indexed on */
/* the GPU (psivox_elem_index_*)                                       */
/* ------------------------------------------------------------------ */

typedef struct {
    PyObject_HEAD psivox_elem_index* idx;
    PyObject* mesh; /* the Mesh it was built from */
} RStarTree;

/* an element index of mesh on the GPU with the given window */
static int mesh_index(Mesh* mesh, const double win[2][3], psivox_elem_index** out, const char* who) {
    *out = NULL;
    if (mesh->pos == Py_None) {
        PyErr_Format(PyExc_ValueError, "%s: the mesh holds no particles", who);
        return -1;
    }
    if (psivox_device_count() <= 0) {
        PyErr_Format(PyExc_RuntimeError, "%s: no GPU visible to the MI355X voxelizer", who);
        return -1;
    }
    int ret = -1;
    PyObject* pos = contiguous(mesh->pos);
    PyObject* vel = contiguous(mesh->vel);
    PyObject* mass = contiguous(mesh->mass);
    PyObject* conn = contiguous(mesh->connectivity);
    if (pos && vel && mass && conn) {
        psivox_mesh cm;
        memset(&cm, 0, sizeof cm);
        cm.elemtype = mesh->elemtype;
        cm.periodic = mesh->periodic;
        cm.npart = mesh->npart;
        cm.nelem = mesh->nelem;
        cm.pos = (const double*)PyArray_DATA((PyArrayObject*)pos);
        cm.vel = (const double*)PyArray_DATA((PyArrayObject*)vel);
        cm.mass = (const double*)PyArray_DATA((PyArrayObject*)mass);
        cm.conn = (const int32_t*)PyArray_DATA((PyArrayObject*)conn);
        memcpy(cm.box, mesh->box, sizeof cm.box);
        int rc;
        Py_BEGIN_ALLOW_THREADS;
        rc = psivox_elem_index_build(&cm, win, NULL, out);
        Py_END_ALLOW_THREADS;
        if (rc != 0)
            PyErr_Format(rc == PSIVOX_E_NOMEM ? PyExc_MemoryError : PyExc_RuntimeError, "%s: %s", who,
                         psivox_last_error());
        else
            ret = 0;
    }
    Py_XDECREF(pos);
    Py_XDECREF(vel);
    Py_XDECREF(mass);
    Py_XDECREF(conn);
    return ret;
}

static int RStarTree_init(RStarTree* self, PyObject* args, PyObject* kwds) {
    Mesh* mesh = NULL;
    int init_cap = 1024; /* accepted and unused, as in the reference (src/PSI.c:889) */
    static char* kwlist[] = {"mesh", "initial_capacity", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!|i", kwlist, &MeshType, &mesh, &init_cap)) return -1;
    (void)init_cap;
    /* the periodic box, or an unbounded window (src/PSI.c:875-886) */
    double win[2][3];
    for (int a = 0; a < 3; ++a) {
        win[0][a] = mesh->periodic ? mesh->box[0][a] : -1.0e30;
        win[1][a] = mesh->periodic ? mesh->box[1][a] : 1.0e30;
    }
    psivox_elem_index* idx;
    if (mesh_index(mesh, win, &idx, "PSI.RStarTree()") < 0) return -1;
    psivox_elem_index_free(self->idx);
    self->idx = idx;
    Py_INCREF(mesh);
    Py_XSETREF(self->mesh, (PyObject*)mesh);
    return 0;
}

static PyObject* RStarTree_new(PyTypeObject* type, PyObject* args, PyObject* kwds) {
    (void)args;
    (void)kwds;
    RStarTree* self = (RStarTree*)type->tp_alloc(type, 0);
    if (self) {
        self->idx = NULL;
        self->mesh = NULL;
    }
    return (PyObject*)self;
}

static void RStarTree_dealloc(RStarTree* self) {
    psivox_elem_index_free(self->idx);
    Py_XDECREF(self->mesh);
    Py_TYPE(self)->tp_free((PyObject*)self);
}

/* query(box): the elements whose box meets the closed box, as an int32
 * array in ascending order (src/PSI.c:817-861 returns the same elements in
 * its tree's order) */
static PyObject* RStarTree_query(RStarTree* self, PyObject* args, PyObject* kwds) {
    PyObject* box;
    double q[2][3];
    static char* kwlist[] = {"box", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O", kwlist, &box)) return NULL;
    if (!PyArg_ParseTuple(box, "(ddd)(ddd)", &q[0][0], &q[0][1], &q[0][2], &q[1][0], &q[1][1], &q[1][2])) {
        PyErr_SetString(PyExc_TypeError, "Query box must be a tuple containing two tuples of 3 values each");
        return NULL;
    }
    if (!self->idx) {
        PyErr_SetString(PyExc_RuntimeError, "RStarTree was not initialised");
        return NULL;
    }
    int64_t cap = 1024, n = 0;
    for (;;) {
        npy_intp d = (npy_intp)cap;
        PyObject* out = PyArray_SimpleNew(1, &d, NPY_INT32);
        if (!out) return NULL;
        int rc;
        int32_t* od = (int32_t*)PyArray_DATA((PyArrayObject*)out);
        Py_BEGIN_ALLOW_THREADS;
        rc = psivox_elem_index_query(self->idx, (const double(*)[3])q, od, cap, &n, NULL);
        Py_END_ALLOW_THREADS;
        if (rc != 0) {
            Py_DECREF(out);
            PyErr_Format(PyExc_RuntimeError, "RStarTree.query(): %s", psivox_last_error());
            return NULL;
        }
        if (n <= cap) {
            PyArray_Dims nd = {&d, 1};
            d = (npy_intp)n;
            PyObject* r = PyArray_Resize((PyArrayObject*)out, &nd, 0, NPY_CORDER);
            if (!r) {
                Py_DECREF(out);
                return NULL;
            }
            Py_DECREF(r);
            return out;
        }
        Py_DECREF(out);
        cap = n;
    }
}

static PyMethodDef RStarTree_methods[] = {
    {"query", (PyCFunction)(void (*)(void))RStarTree_query, METH_VARARGS | METH_KEYWORDS, "Query the RStarTree."},
    {NULL, NULL, 0, NULL}};

static PyTypeObject RStarTreeType = {
    PyVarObject_HEAD_INIT(NULL, 0).tp_name = "PSI.RStarTree",
    .tp_basicsize = sizeof(RStarTree),
    .tp_dealloc = (destructor)RStarTree_dealloc,
    .tp_flags = Py_TPFLAGS_DEFAULT | Py_TPFLAGS_BASETYPE,
    .tp_doc = "Element bounding boxes of a Mesh, indexed on the GPU",
    .tp_methods = RStarTree_methods,
    .tp_init = (initproc)RStarTree_init,
    .tp_new = RStarTree_new,
};

/* VDF(mesh, sample_pos, tree=None, refine_tolerance=1.0, refine_max_lvl=0)
 * -> (rho, vel): the streams at sample_pos (src/PSI.c:1022-1058,
 * src/cpsi.c:177-275).  Without a tree, the elements are indexed with the
 * mesh's box as the window, as psi_sample_vdf does (src/cpsi.c:203-208;
 * for a non-periodic mesh that box is all zeros, so only elements whose box
 * holds the origin take part there too). */
static PyObject* PSI_VDF(PyObject* self, PyObject* args, PyObject* kwds) {
    (void)self;
    Mesh* mesh = NULL;
    RStarTree* tree = NULL;
    double x[3], reftol = 1.0;
    int maxlvl = 0;
    static char* kwlist[] = {"mesh", "sample_pos", "tree", "refine_tolerance", "refine_max_lvl", NULL};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!(ddd)|O!di", kwlist, &MeshType, &mesh, &x[0], &x[1], &x[2],
                                     &RStarTreeType, &tree, &reftol, &maxlvl))
        return NULL;
    psivox_elem_index* idx = NULL;
    int own = 0;
    if (tree) {
        if (!tree->idx) {
            PyErr_SetString(PyExc_RuntimeError, "RStarTree was not initialised");
            return NULL;
        }
        if (tree->mesh != (PyObject*)mesh) {
            PyErr_SetString(PyExc_ValueError, "PSI.VDF(): tree was built from a different mesh");
            return NULL;
        }
        idx = tree->idx;
    } else {
        if (mesh_index(mesh, (const double(*)[3])mesh->box, &idx, "PSI.VDF()") < 0) return NULL;
        own = 1;
    }
    PyObject *rho = NULL, *vel = NULL, *ret = NULL;
    int64_t cap = 256, n = 0;
    for (;;) {
        npy_intp d[2] = {(npy_intp)cap, 3};
        rho = PyArray_SimpleNew(1, d, NPY_FLOAT64);
        vel = PyArray_SimpleNew(2, d, NPY_FLOAT64);
        if (!rho || !vel) goto done;
        int rc;
        double* rd = (double*)PyArray_DATA((PyArrayObject*)rho);
        double* vd = (double*)PyArray_DATA((PyArrayObject*)vel);
        Py_BEGIN_ALLOW_THREADS;
        rc = psivox_sample_vdf(idx, x, reftol, maxlvl, rd, vd, cap, &n, NULL);
        Py_END_ALLOW_THREADS;
        if (rc != 0) {
            PyErr_Format(rc == PSIVOX_E_NOMEM ? PyExc_MemoryError : PyExc_RuntimeError, "PSI.VDF(): %s",
                         psivox_last_error());
            goto done;
        }
        if (n <= cap) break;
        Py_CLEAR(rho);
        Py_CLEAR(vel);
        cap = n;
    }
    {
        npy_intp d[2] = {(npy_intp)n, 3};
        PyArray_Dims d1 = {d, 1}, d2 = {d, 2};
        PyObject* r1 = PyArray_Resize((PyArrayObject*)rho, &d1, 0, NPY_CORDER);
        if (!r1) goto done;
        Py_DECREF(r1);
        PyObject* r2 = PyArray_Resize((PyArrayObject*)vel, &d2, 0, NPY_CORDER);
        if (!r2) goto done;
        Py_DECREF(r2);
    }
    ret = Py_BuildValue("OO", rho, vel);
done:
    Py_XDECREF(rho);
    Py_XDECREF(vel);
    if (own) psivox_elem_index_free(idx);
    return ret;
}

static PyMethodDef module_methods[] = {
    {"voxels", (PyCFunction)(void (*)(void))PSI_voxels, METH_VARARGS | METH_KEYWORDS, "Voxelizes"},
    {"VDF", (PyCFunction)(void (*)(void))PSI_VDF, METH_VARARGS | METH_KEYWORDS, "VDF"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef PSImod = {PyModuleDef_HEAD_INIT, "PSI",
                                    "PSI, the phase space intersector (MI355X voxelizer)", -1, module_methods};

PyMODINIT_FUNC PyInit_PSI(void) {
    PyObject* m = PyModule_Create(&PSImod);
    if (!m) return NULL;
    import_array();
    if (PyType_Ready(&MeshType) < 0) return NULL;
    Py_INCREF(&MeshType);
    PyModule_AddObject(m, "Mesh", (PyObject*)&MeshType);
    if (PyType_Ready(&GridType) < 0) return NULL;
    Py_INCREF(&GridType);
    PyModule_AddObject(m, "Grid", (PyObject*)&GridType);
    if (PyType_Ready(&RStarTreeType) < 0) return NULL;
    Py_INCREF(&RStarTreeType);
    PyModule_AddObject(m, "RStarTree", (PyObject*)&RStarTreeType);
    return m;
}
