// Status codes of the eithot C-ABI (shared by the host planner and runtime).
#pragma once
#define EITHOT_OK 0
#define EITHOT_ERR_INVALID (-1)      /* bad dims, permutation, order or alpha */
#define EITHOT_ERR_ELEM (-2)         /* unsupported element size */
#define EITHOT_ERR_ALPHA (-3)        /* the plan needs more scratch than alpha allows */
#define EITHOT_ERR_DEVICE (-4)       /* HIP error; tensor contents are indeterminate */
#define EITHOT_ERR_NOMEM (-5)        /* scratch allocation failed; tensor untouched */
#define EITHOT_ERR_UNSUPPORTED (-6)  /* operation not available */
#define EITHOT_ERR_NO_DEVICE (-7)    /* no HIP device present */
#define EITHOT_ERR_BUSY (-8)         /* serving: request refused by admission */
