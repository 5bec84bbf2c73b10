// Module initialisation for a reduced build of the reference C extension.
//
// TEST INFRASTRUCTURE ONLY.  This file is linked with the reference's own
// sources for the FD grid path (c/bmgs/bmgs.cpp, c/bc.cpp, c/operators.cpp,
// c/transformers.cpp, c/fd_preconditioner.cpp, c/woperators.cpp), compiled
// where they lie under the reference tree by oracle/Makefile.  It registers
// the constructors the reference method table exposes for this path
// (c/_gpaw.h:182,183,185,194)
// under a module named `_gpaw_ref`, so golden vectors can be produced from the
// real reference code in the build container.  Nothing in the product links,
// imports or calls this module.

#define __GPAW_SHOULD_IMPORT_NUMPY
#include "python_utils.h"

extern "C++" {
PyObject* NewOperatorObject(PyObject* self, PyObject* args);
PyObject* NewTransformerObject(PyObject* self, PyObject* args);
PyObject* fd_precond(PyObject* self, PyObject* args);
PyObject* NewWOperatorObject(PyObject* self, PyObject* args);
extern PyTypeObject OperatorType;
extern PyTypeObject TransformerType;
extern PyTypeObject WOperatorType;
}

static PyMethodDef ref_functions[] = {
    {"Operator", NewOperatorObject, METH_VARARGS, 0},
    {"Transformer", NewTransformerObject, METH_VARARGS, 0},
    {"fd_precond", fd_precond, METH_VARARGS, 0},
    {"WOperator", NewWOperatorObject, METH_VARARGS, 0},
    {0, 0, 0, 0}};

static struct PyModuleDef ref_module = {
    PyModuleDef_HEAD_INIT, "_gpaw_ref",
    "Reference FD-path objects (operators, transformers, fd_precond)", -1,
    ref_functions, 0, 0, 0, 0};

PyMODINIT_FUNC PyInit__gpaw_ref(void)
{
    import_array1(0);
    if (PyType_Ready(&OperatorType) < 0) return 0;
    if (PyType_Ready(&TransformerType) < 0) return 0;
    if (PyType_Ready(&WOperatorType) < 0) return 0;
    PyObject* m = PyModule_Create(&ref_module);
    if (m == 0) return 0;
    Py_INCREF(&OperatorType);
    Py_INCREF(&TransformerType);
    Py_INCREF(&WOperatorType);
    return m;
}
