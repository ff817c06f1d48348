/* dqsim_opstream.h -- native opcode-stream wire format for circuits.
 *
 * An alternative to the circuit JSON of src/types.rs:7-135 (serde,
 * `kind`-tagged instructions) for large circuits: an array of fixed-size
 * little-endian records that a host writes directly, with no JSON encoding
 * or parsing.  Each record is one instruction of the JSON schema, with the
 * same qubit order (qubits[0] = first qubit field = matrix MSB) and the
 * parameters in the JSON field order.  A `conditional` record is followed by
 * the record of its instruction.  Generic `gate` names live in a separate
 * string table; `cbit` holds the table index.  No reference counterpart.
 */
#ifndef DQSIM_OPSTREAM_H
#define DQSIM_OPSTREAM_H

#include <stdint.h>

enum dq_opcode {
    DQ_OP_ID = 0, DQ_OP_X, DQ_OP_Y, DQ_OP_Z, DQ_OP_H, DQ_OP_S, DQ_OP_SDG, DQ_OP_T, DQ_OP_TDG,
    DQ_OP_SX, DQ_OP_SXDG, DQ_OP_U0,
    DQ_OP_U3, DQ_OP_U, DQ_OP_U2, DQ_OP_U1, DQ_OP_P, DQ_OP_RX, DQ_OP_RY, DQ_OP_RZ,
    DQ_OP_CX, DQ_OP_CZ, DQ_OP_CY, DQ_OP_CH, DQ_OP_CSX, DQ_OP_SWAP,
    DQ_OP_CRX, DQ_OP_CRY, DQ_OP_CRZ, DQ_OP_CU1, DQ_OP_CP, DQ_OP_CU3, DQ_OP_CU, DQ_OP_RXX, DQ_OP_RZZ,
    DQ_OP_CCX, DQ_OP_CSWAP, DQ_OP_RCCX, DQ_OP_RC3X, DQ_OP_C3X, DQ_OP_C3SQRTX, DQ_OP_C4X,
    DQ_OP_GATE,        /* generic gate: name = names[cbit], qubits[0..nq) */
    DQ_OP_MEASURE,     /* qubits[0] -> cbit */
    DQ_OP_RESET,       /* qubits[0] */
    DQ_OP_CONDITIONAL, /* qubits[0] = creg_base, qubits[1] = creg_size, value = creg_value; next record = op */
    DQ_OP_BARRIER,
    DQ_OP_CLASSICAL,   /* name = names[cbit] */
    DQ_OP_COUNT
};

typedef struct dq_op_record {
    uint16_t opcode;     /* enum dq_opcode */
    uint16_t nq;         /* qubits used (gate: its qubit count) */
    uint32_t cbit;       /* measure: cbit; gate / classical: name index */
    uint32_t qubits[5];
    uint32_t pad;
    union {
        double params[4];  /* in the JSON field order of the kind */
        uint64_t value;    /* conditional: creg_value */
    };
} dq_op_record;          /* 64 bytes */

#endif /* DQSIM_OPSTREAM_H */
