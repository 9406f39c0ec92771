/* Python front end helper: packs a list of QPs (dicts with scipy CSC P and A
 * and numpy q, l, u, optional x0, y0) into the arguments of
 * osqp_batch_add_many without copying any array, reading each array's data
 * through the buffer protocol.
 *
 * csc_batch(problems) -> (buf, address) or None
 *   buf: bytes laid out as
 *     [k OSQPCscMatrix P][k OSQPCscMatrix A]
 *     [k pointers to P structs][k q][k pointers to A structs][k l][k u][k x0][k y0]
 *     [k int32 n][k int32 m]
 *   address: the address of buf's first byte (the structs and pointer
 *   tables point into buf, which the caller keeps alive for the call).
 *   None when any problem is not in that exact form (CSC with int32 indices
 *   and float64 values, C-contiguous float64 vectors of lengths n, m, m and
 *   n, m for x0, y0); the caller then converts the arrays itself.
 * The CSC columns must hold ascending row indices (as the engine expects;
 * duplicates allowed) and indptr must not decrease: a matrix that is not
 * so gives None too. */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <stdint.h>
#include <string.h>

#include "osqp_amd.h"

typedef struct {
  const void* ptr;
  Py_ssize_t len;
} Arr;

/* a C-contiguous 1-D buffer of itemsize bytes whose format is one of fmt's
 * accepted codes; the buffer is released at once (the caller holds the
 * object for the duration of the add) */
static int get_arr(PyObject* o, Py_ssize_t itemsize, char code, Arr* out) {
  Py_buffer v;
  if (PyObject_GetBuffer(o, &v, PyBUF_C_CONTIGUOUS | PyBUF_FORMAT) != 0) {
    PyErr_Clear();
    return 0;
  }
  int good = v.itemsize == itemsize && v.ndim <= 1 && v.format != NULL;
  if (good) {
    const char* f = v.format;
    if (*f == '<' || *f == '=' || *f == '@') f++;
    good = f[0] == code && f[1] == '\0';
  }
  out->ptr = v.buf;
  out->len = v.itemsize ? v.len / v.itemsize : 0;
  PyBuffer_Release(&v);
  return good;
}

static int get_vec(PyObject* d, const char* key, Py_ssize_t n, int optional, Arr* out) {
  PyObject* o = PyDict_GetItemString(d, key); /* borrowed */
  if (o == NULL || o == Py_None) {
    out->ptr = NULL;
    out->len = 0;
    return optional;
  }
  return get_arr(o, 8, 'd', out) && out->len == n;
}

/* #{k : a[k + 1] < a[k]}: 32-bit compares counted in 32 bits, so the loop
 * vectorizes */
static __attribute__((noinline)) int32_t count_descents(const int32_t* a, int32_t n) {
  int32_t d = 0;
  for (int32_t k = 0; k < n - 1; ++k) d += a[k + 1] < a[k];
  return d;
}

/* indptr non-decreasing, and the row indices ascending within each column:
 * every descent i[k] < i[k - 1] must sit at a column start */
static int sorted_csc(const int32_t* pp, const int32_t* ii, long cols) {
  for (long j = 0; j < cols; ++j)
    if (pp[j + 1] < pp[j]) return 0;
  const int32_t nnz = pp[cols];
  const int32_t desc = count_descents(ii, nnz);
  if (desc == 0) return 1;
  int32_t at_starts = 0;
  int32_t last = 0;
  for (long j = 1; j < cols; ++j) {
    const int32_t k = pp[j];
    if (k > last && k < nnz) {
      at_starts += ii[k] < ii[k - 1];
      last = k;
    }
  }
  return at_starts == desc;
}

/* P or A: format "csc", shape (rows, cols), indptr int32 of cols + 1 with
 * indptr[0] = 0, indices int32 and data float64 of at least nnz entries,
 * columns sorted */
static int get_csc(PyObject* d, const char* key, OSQPCscMatrix* M) {
  PyObject* o = PyDict_GetItemString(d, key);
  if (o == NULL) return 0;
  int good = 0;
  PyObject *fmt = NULL, *shape = NULL, *ip = NULL, *ix = NULL, *dt = NULL;
  fmt = PyObject_GetAttrString(o, "format");
  if (fmt == NULL || !PyUnicode_Check(fmt) || PyUnicode_CompareWithASCIIString(fmt, "csc") != 0) goto done;
  shape = PyObject_GetAttrString(o, "shape");
  if (shape == NULL || !PyTuple_Check(shape) || PyTuple_GET_SIZE(shape) != 2) goto done;
  {
    const long rows = PyLong_AsLong(PyTuple_GET_ITEM(shape, 0));
    const long cols = PyLong_AsLong(PyTuple_GET_ITEM(shape, 1));
    if (PyErr_Occurred() || rows < 0 || cols < 0 || rows > INT32_MAX || cols >= INT32_MAX) goto done;
    ip = PyObject_GetAttrString(o, "indptr");
    ix = PyObject_GetAttrString(o, "indices");
    dt = PyObject_GetAttrString(o, "data");
    if (ip == NULL || ix == NULL || dt == NULL) goto done;
    Arr a_p, a_i, a_x;
    if (!get_arr(ip, 4, 'i', &a_p) || a_p.len != cols + 1) goto done;
    if (!get_arr(ix, 4, 'i', &a_i) || !get_arr(dt, 8, 'd', &a_x)) goto done;
    const int32_t* pp = (const int32_t*)a_p.ptr;
    const int32_t nnz = pp[cols];
    if (pp[0] != 0 || nnz < 0 || a_i.len < nnz || a_x.len < nnz) goto done;
    if (!sorted_csc(pp, (const int32_t*)a_i.ptr, cols)) goto done;
    M->m = (OSQPInt)rows;
    M->n = (OSQPInt)cols;
    M->p = (OSQPInt*)pp;
    M->i = (OSQPInt*)a_i.ptr;
    M->x = (OSQPFloat*)a_x.ptr;
    M->nzmax = nnz;
    M->nz = -1;
    M->owned = 0;
    good = 1;
  }
done:
  PyErr_Clear();
  Py_XDECREF(fmt);
  Py_XDECREF(shape);
  Py_XDECREF(ip);
  Py_XDECREF(ix);
  Py_XDECREF(dt);
  return good;
}

static PyObject* csc_batch(PyObject* self, PyObject* args) {
  (void)self;
  PyObject* seq;
  if (!PyArg_ParseTuple(args, "O", &seq)) return NULL;
  PyObject* fast = PySequence_Fast(seq, "problems must be a sequence");
  if (fast == NULL) return NULL;
  const Py_ssize_t k = PySequence_Fast_GET_SIZE(fast);
  const size_t s_mat = sizeof(OSQPCscMatrix), s_ptr = sizeof(void*);
  const size_t off_A = (size_t)k * s_mat, off_tab = 2 * (size_t)k * s_mat;
  const size_t off_dims = off_tab + 7 * (size_t)k * s_ptr;
  PyObject* buf = PyBytes_FromStringAndSize(NULL, (Py_ssize_t)(off_dims + 2 * (size_t)k * sizeof(int32_t)));
  if (buf == NULL) {
    Py_DECREF(fast);
    return NULL;
  }
  char* b = PyBytes_AS_STRING(buf);
  OSQPCscMatrix* Ps = (OSQPCscMatrix*)b;
  OSQPCscMatrix* As = (OSQPCscMatrix*)(b + off_A);
  const void** tab = (const void**)(b + off_tab); /* P q A l u x0 y0, k each */
  int32_t* ns = (int32_t*)(b + off_dims);
  int32_t* ms = ns + k;
  int good = 1;
  for (Py_ssize_t j = 0; j < k && good; ++j) {
    PyObject* d = PySequence_Fast_GET_ITEM(fast, j);
    if (!PyDict_Check(d) || !get_csc(d, "P", &Ps[j]) || !get_csc(d, "A", &As[j])) {
      good = 0;
      break;
    }
    const Py_ssize_t n = Ps[j].n, m = As[j].m;
    Arr q, l, u, x0, y0;
    good = Ps[j].m == n && As[j].n == n && get_vec(d, "q", n, 0, &q) && get_vec(d, "l", m, 0, &l) &&
           get_vec(d, "u", m, 0, &u) && get_vec(d, "x0", n, 1, &x0) && get_vec(d, "y0", m, 1, &y0);
    if (!good) break;
    tab[0 * k + j] = &Ps[j];
    tab[1 * k + j] = q.ptr;
    tab[2 * k + j] = &As[j];
    tab[3 * k + j] = l.ptr;
    tab[4 * k + j] = u.ptr;
    tab[5 * k + j] = x0.ptr;
    tab[6 * k + j] = y0.ptr;
    ns[j] = (int32_t)n;
    ms[j] = (int32_t)m;
  }
  Py_DECREF(fast);
  if (!good) {
    Py_DECREF(buf);
    Py_RETURN_NONE;
  }
  PyObject* addr = PyLong_FromVoidPtr(b);
  if (addr == NULL) {
    Py_DECREF(buf);
    return NULL;
  }
  PyObject* out = PyTuple_Pack(2, buf, addr);
  Py_DECREF(buf);
  Py_DECREF(addr);
  return out;
}

static PyMethodDef methods[] = {
    {"csc_batch", csc_batch, METH_VARARGS, "Pack QPs into osqp_batch_add_many's arguments (or None)."},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef module = {PyModuleDef_HEAD_INIT, "_marshal", NULL, -1, methods, NULL, NULL, NULL, NULL};

PyMODINIT_FUNC PyInit__marshal(void) { return PyModule_Create(&module); }
