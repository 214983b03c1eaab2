/*
 * Private object layouts of the host core (the libCEED API implementation).
 * Each object carries the backend slots CeedSetBackendFunction fills, a
 * reference count and the backend's data pointer.
 */
#ifndef CEED_CORE_H
#define CEED_CORE_H

#include <ceed.h>
#include <ceed/backend.h>
#include <stdint.h>

#define CEED_CORE_MAX_FIELDS CEED_FIELD_MAX

struct Ceed_private {
  char *resource;
  int   ref_count;
  bool  is_deterministic;
  void *data;
  CeedErrorHandler Error;
  char             err_msg[CEED_MAX_RESOURCE_LEN];
  /* backend slots */
  int (*GetPreferredMemType)(CeedMemType *);
  int (*Destroy)(Ceed);
  int (*VectorCreate)(CeedSize, CeedVector);
  int (*ElemRestrictionCreate)(CeedMemType, CeedCopyMode, const CeedInt *, const bool *, const CeedInt8 *, CeedElemRestriction);
  int (*BasisCreateTensorH1)(CeedInt, CeedInt, CeedInt, const CeedScalar *, const CeedScalar *, const CeedScalar *, const CeedScalar *, CeedBasis);
  int (*QFunctionCreate)(CeedQFunction);
  int (*QFunctionContextCreate)(CeedQFunctionContext);
  int (*OperatorCreate)(CeedOperator);
};

struct CeedVector_private {
  Ceed     ceed;
  int      ref_count;
  CeedSize length;
  uint64_t state;      /* odd while a write access is out */
  int      num_readers;
  void    *data;
  int (*HasValidArray)(CeedVector, bool *);
  int (*HasBorrowedArrayOfType)(CeedVector, CeedMemType, bool *);
  int (*SetArray)(CeedVector, CeedMemType, CeedCopyMode, CeedScalar *);
  int (*TakeArray)(CeedVector, CeedMemType, CeedScalar **);
  int (*SyncArray)(CeedVector, CeedMemType);
  int (*SetValue)(CeedVector, CeedScalar);
  int (*GetArray)(CeedVector, CeedMemType, CeedScalar **);
  int (*GetArrayRead)(CeedVector, CeedMemType, const CeedScalar **);
  int (*GetArrayWrite)(CeedVector, CeedMemType, CeedScalar **);
  int (*RestoreArray)(CeedVector);
  int (*RestoreArrayRead)(CeedVector);
  int (*Norm)(CeedVector, CeedNormType, CeedScalar *);
  int (*Scale)(CeedVector, CeedScalar);
  int (*AXPY)(CeedVector, CeedScalar, CeedVector);
  int (*AXPBY)(CeedVector, CeedScalar, CeedScalar, CeedVector);
  int (*PointwiseMult)(CeedVector, CeedVector, CeedVector);
  int (*Reciprocal)(CeedVector);
  int (*Destroy)(CeedVector);
};

struct CeedElemRestriction_private {
  Ceed                ceed;
  int                 ref_count;
  CeedRestrictionType type;
  CeedInt             num_elem, elem_size, num_comp, comp_stride;
  CeedSize            l_size;
  CeedInt             strides[3];
  bool                backend_strides;
  CeedInt             e_layout[3];
  uint64_t            num_readers; /* offsets arrays handed out and not yet restored */
  void               *data;
  int (*Apply)(CeedElemRestriction, CeedTransposeMode, CeedVector, CeedVector, CeedRequest *);
  int (*GetOffsets)(CeedElemRestriction, CeedMemType, const CeedInt **);
  int (*Destroy)(CeedElemRestriction);
};

struct CeedBasis_private {
  Ceed        ceed;
  int         ref_count;
  CeedInt     dim, num_comp, P_1d, Q_1d;
  CeedScalar *interp_1d, *grad_1d, *q_ref_1d, *q_weight_1d;
  void       *data;
  int (*Apply)(CeedBasis, CeedInt, CeedTransposeMode, CeedEvalMode, CeedVector, CeedVector);
  int (*ApplyAdd)(CeedBasis, CeedInt, CeedTransposeMode, CeedEvalMode, CeedVector, CeedVector);
  int (*Destroy)(CeedBasis);
};

struct CeedQFunctionField_private {
  char        *name;
  CeedInt      size;
  CeedEvalMode eval_mode;
};

struct CeedQFunctionContext_private {
  Ceed     ceed;
  int      ref_count;
  size_t   size;
  uint64_t state, num_readers;
  void    *backend_data; /* the backend stores the data itself */
  int (*HasValidData)(CeedQFunctionContext, bool *);
  int (*HasBorrowedDataOfType)(CeedQFunctionContext, CeedMemType, bool *);
  int (*SetData)(CeedQFunctionContext, CeedMemType, CeedCopyMode, void *);
  int (*TakeData)(CeedQFunctionContext, CeedMemType, void *);
  int (*GetData)(CeedQFunctionContext, CeedMemType, void *);
  int (*GetDataRead)(CeedQFunctionContext, CeedMemType, void *);
  int (*RestoreData)(CeedQFunctionContext);
  int (*RestoreDataRead)(CeedQFunctionContext);
  int (*Destroy)(CeedQFunctionContext);
};

struct CeedQFunction_private {
  Ceed                              ceed;
  int                               ref_count;
  CeedInt                           vec_length;
  CeedQFunctionUser                 function;
  char                             *source_path, *kernel_name;
  struct CeedQFunctionField_private in[CEED_CORE_MAX_FIELDS], out[CEED_CORE_MAX_FIELDS];
  CeedQFunctionField                in_ptr[CEED_CORE_MAX_FIELDS], out_ptr[CEED_CORE_MAX_FIELDS];
  CeedInt                           num_in, num_out;
  CeedQFunctionContext              ctx;
  CeedSize                          user_flops;
  bool                              is_immutable;
  void                             *data;
  int (*Apply)(CeedQFunction, CeedInt, CeedVector *, CeedVector *);
  int (*Destroy)(CeedQFunction);
};

struct CeedOperatorField_private {
  char               *name;
  CeedElemRestriction rstr;
  CeedBasis           basis;
  CeedVector          vec;
};

struct CeedOperator_private {
  Ceed                             ceed;
  int                              ref_count;
  CeedQFunction                    qf;
  struct CeedOperatorField_private in[CEED_CORE_MAX_FIELDS], out[CEED_CORE_MAX_FIELDS];
  CeedOperatorField                in_ptr[CEED_CORE_MAX_FIELDS], out_ptr[CEED_CORE_MAX_FIELDS];
  CeedInt                          num_fields_set;
  CeedInt                          num_elem, num_qpts;
  bool                             has_num_elem, has_num_qpts, ready;
  bool                             is_composite;
  CeedOperator                    *sub;  /* composite: sub-operators */
  CeedInt                          num_sub;
  void                            *data;
  int (*Apply)(CeedOperator, CeedVector, CeedVector, CeedRequest *);
  int (*ApplyAdd)(CeedOperator, CeedVector, CeedVector, CeedRequest *);
  int (*LinearAssembleDiagonal)(CeedOperator, CeedVector, CeedRequest *);
  int (*LinearAssembleAddDiagonal)(CeedOperator, CeedVector, CeedRequest *);
  int (*Destroy)(CeedOperator);
};

/* core-internal helpers */
int CeedCoreIsSpecialVector(CeedVector v);
int CeedCoreGalleryLookup(const char *name, CeedQFunction qf);
int CeedCoreJitRoot(const char **root);

#endif /* CEED_CORE_H */
